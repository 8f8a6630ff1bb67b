/*
 * roift.h -- C-ABI of the MI355X (gfx950) OIFT f_max + fuzzy relaxation library
 * (libroift.so).  Plain pointers and sizes only; every call returns a status.
 *
 * Conventions (same as the gft library the reference builds on):
 *   - volumes are x-fastest: voxel p = x + y*xsize + z*xsize*ysize
 *     (gft_scene32.h:173-179);
 *   - labels: -1 (NIL) unlabelled, 0 background, 1..L objects;
 *   - seeds: the gft S array without its count word, i.e. S[1..S[0]] in
 *     insertion order (gft_ift.cpp:938-943); the order is the FIFO tie order;
 *   - adjacency: gft::AdjRel3::Spheric(r) (gft_adjrel3.cpp:47-85); r = 1.0
 *     gives 6-adjacency (a 2-D image is zsize = 1), r = 1.75 gives 26;
 *   - polarity "per" is alpha x 100 in [-100, 100]; per_class[c] likewise.
 *
 * Pointers named d_* are device (HIP) pointers on the context's device; h_*
 * are host pointers.  Calls on one context are serialised on its stream and
 * return after the device work is complete.  A context is not thread-safe.
 */
#ifndef ROIFT_H
#define ROIFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ROIFT_API __attribute__((visibility("default")))

enum roift_status {
    ROIFT_OK = 0,
    ROIFT_EINVAL = -1,      /* bad argument (shape, seed index, NIL seed label, duplicate seed, negative image) */
    ROIFT_ENOMEM = -2,      /* device or host allocation failed */
    ROIFT_EHIP = -3,        /* a HIP runtime call failed */
    ROIFT_EUNSUPPORTED = -4 /* valid input outside what this build implements (see roift_last_error) */
};

enum roift_dtype { ROIFT_I16 = 1, ROIFT_I32 = 2, ROIFT_U16 = 3, ROIFT_U32 = 4, ROIFT_F32 = 5, ROIFT_U8 = 6, ROIFT_I8 = 7 };

typedef struct roift_ctx roift_ctx;

/* Problem description shared by the entry points. */
typedef struct {
    int xsize, ysize, zsize;
    float adj_radius;          /* Spheric radius: 1.0 (6-adj) or 1.75 (26-adj) */
    int use_per_class;         /* 0: OIFT_Multi / ORelax_1_Multi polarity from "per"
                                  (label 0 -> -per, others +per);
                                  1: *_PerClass polarity per_class[label] */
    float per;                 /* polarity x100 when use_per_class == 0 */
    const float *h_per_class;  /* length max_label+1 when use_per_class == 1 */
    int max_label;
    float dx, dy, dz;          /* voxel spacing (gft's scn->dx/dy/dz); 0 reads as 1.0, so a
                                  zero-initialised tail means unit spacing; any other finite
                                  value is used as gft uses it.  Read by the
                                  relaxations' Dpq = sqrtf(sum (d_i spacing_i)^2) / min spacing
                                  (gft_ift.cpp:8723-8731) and the Gaussian blur's taps
                                  (Kernel3::SphericalGaussian(scn, ...), gft_filtering3.cpp:106-126);
                                  the OIFT's arc weights do not depend on it. */
} roift_problem;

/* ---- context ---------------------------------------------------------------- */
ROIFT_API int roift_create(roift_ctx **ctx, int device);
ROIFT_API void roift_destroy(roift_ctx *ctx);
ROIFT_API const char *roift_last_error(const roift_ctx *ctx);
/* Run on this hipStream_t (default: a stream the context creates). */
ROIFT_API int roift_set_stream(roift_ctx *ctx, void *hip_stream);
/* Stage timings (device events) of the last call, in milliseconds:
   [0] weights, [1] oift, [2] relax, [3] labels; [4] OIFT level steps, [5] flood rounds. */
ROIFT_API int roift_last_timings(const roift_ctx *ctx, double *out6);
/* Device-event times of the last call's main kernels, in milliseconds:
   [0] weight builder, [1] OIFT persistent kernel, [2] all relaxation iterations
   (back-to-back launches), [3] number of relaxation launches; with n >= 5,
   [4] the last Gaussian blur (all passes); with n >= 6, [5] the last Sobel /
   distance / boundary-seed / dilation call (from its first kernel to its last).  n >= 4.
   (No reference counterpart: measurement only.) */
ROIFT_API int roift_kernel_timings(const roift_ctx *ctx, double *out, int n);
/* Wall-clock breakdown of the last OIFT's persistent kernel, measured on the
   device between its grid barriers: out[0..8] milliseconds spent in the phases
   level-scan, collect, key-collect, key-rank, flood-seed, flood-round,
   r_safe, claim, offers; out[9..17] how many times each phase ran; with
   n >= 162, out[18 + 16*phase + b] splits each phase's time by the step's
   entry count (bucket b = ceil(log2(entries)), capped at 15); with n >= 164,
   out[162] = items the flood rounds processed, out[163] = touched-list slots
   the claims scanned; with n >= 166, out[164] / out[165] = flood batches summed
   over / maximal among the workgroups; with n >= 167, out[166] = how many OIFT runs of
   this context in ROIFT_PARAM_OIFT_MODE 1 met a tie and reran level-synchronously; with
   n >= 168, out[167] = the flood's critical path: the sum over flood rounds of the most
   batches one workgroup ran in the round; out[168 + b] (n >= 178) = flood batches of
   ceil(log2(items)) = b, b = 0..9; out[178] (n >= 179) = the items those busiest
   workgroups ran, summed over the rounds; out[179] / out[180] (n >= 181) = the flood hops
   run by a single wavefront on a narrow queue and the ms they took, summed over workgroups.
   After a run in that mode that did not, out[0..4] are instead its grid rounds,
   levels, candidate checks, voxels the checks visited and grid-wide flood rounds, and
   out[5..9] the ms of its phases (frontier, candidates, checks, levels, grid floods) and
   out[10] the most voxels one workgroup's level settled before any grid-wide flood,
   out[11] / out[12] the passes of each round's longest check / the steps of its longest
   level, summed over the rounds.
   n >= 18.                                                                      */
ROIFT_API int roift_oift_profile(const roift_ctx *ctx, double *out, int n);

/* Tuning knobs (the autotuner in roift_amd/tune.py sets them):
   ROIFT_PARAM_OIFT_BLOCKS    workgroups of the persistent OIFT kernel (0 = 0.27 * cbrt(voxels *
                              (adjacency - 1)), a multiple of 8, at least 32, at most the co-resident limit)
   ROIFT_PARAM_FLOOD_BUDGET   LDS batches a workgroup relaxes per flood round before spilling (0 = 1024)
   ROIFT_PARAM_RELAX_DEPTH    reserved: accepted and ignored (each relaxation iteration is one
                              launch; the kernels are bound by arithmetic, not by HBM)
   ROIFT_PARAM_FLOOD_MODE     OIFT flood unit: 0 = single voxels (default; a queue of at most
                              64 items is run by one wavefront without workgroup barriers),
                              1 = 8x8x8 bricks relaxed in LDS (6-adjacency only; slower on thin
                              fronts), 2 = single voxels, every batch workgroup-wide (for measurement)
   ROIFT_PARAM_RELAX_TILE     two-class 6-adjacency relaxation tile: rows (1..8) | planes << 8
                              | xcd << 16 (1: neighbouring tiles on one XCD); 0 = default
   ROIFT_PARAM_BLUR_PLANES    Gaussian blur: z planes per workgroup (0 = automatic)
   ROIFT_PARAM_OIFT_MODE      roift_oift_edges_dev: 0 = level-synchronous (default, exact on every
                              map); 1 = local steps for tie-free maps (pairwise distinct oriented
                              arc weights, e.g. dense ranks): every level that nothing before it
                              can reach runs at once, many levels per grid round.  Exact on
                              tie-free maps; when the run meets a tie it reruns level-synchronously
                              (same result, the time of both).
   ROIFT_PARAM_LOCAL_CANDIDATES  OIFT_MODE 1: pending voxels tested per grid round (0 = 256)
   ROIFT_PARAM_LOCAL_CHECK_STEPS OIFT_MODE 1: passes (512 arcs each) a candidate's test takes before
                              it gives up (0 = 256; the smallest pending voxel runs unchecked)
   ROIFT_PARAM_OIFT_PROFILE   1: the persistent OIFT also gathers the flood's critical path, its
                              batch-size histogram and the single-wave hop times (roift_oift_profile
                              out[167..180]); they cost a few device atomics per flood round, so the
                              default (0) leaves them out and reports them as 0                 */
enum roift_param {
    ROIFT_PARAM_OIFT_BLOCKS = 1,
    ROIFT_PARAM_FLOOD_BUDGET = 2,
    ROIFT_PARAM_RELAX_DEPTH = 3,
    ROIFT_PARAM_FLOOD_MODE = 4,
    ROIFT_PARAM_RELAX_TILE = 5,
    ROIFT_PARAM_BLUR_PLANES = 6,
    ROIFT_PARAM_OIFT_MODE = 7,
    ROIFT_PARAM_LOCAL_CANDIDATES = 8,
    ROIFT_PARAM_LOCAL_CHECK_STEPS = 9,
    ROIFT_PARAM_OIFT_PROFILE = 10
};
ROIFT_API int roift_set_param(roift_ctx *ctx, int param, int value);

/* ---- oriented arc-weight builder --------------------------------------------- */
/* Replaces the weight computed inline in OIFT_Multi (gft_ift.cpp:965-973).
 * For every forward neighbour slot s (adjacency index i = (n-1)/2 + 1 + s,
 * q = p + d_i, E = (n-1)/2 slots) writes
 *     d_wa[s*N + p] = w_obj(p->q),  d_wb[s*N + p] = w_obj(q->p)
 * with w_obj the +per arc weight; the background (-per) weight is the
 * transpose: w_bg(p->q) = w_obj(q->p).  Out-of-volume slots hold the dtype's
 * maximum.  weight_dtype is ROIFT_U16 (requires Wmax <= 65534) or ROIFT_U32.
 * h_maxI / h_minI (nullable) receive the image extrema.                        */
ROIFT_API int roift_weights_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                void *d_wa, void *d_wb, int weight_dtype, int32_t *h_maxI, int32_t *h_minI);

/* roift_weights_dev without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, kernels and results as roift_weights_dev;
 * its scratch (per-workgroup extrema partials and the overflow flag) is a static device
 * array of 8 MiB + 256 B that every GPU the library runs on holds until the process ends,
 * sized for ceil(xsize/64) * ceil(ysize/4) * ceil(zsize/16) <= 2^20 workgroups; larger
 * volumes return ROIFT_EUNSUPPORTED before any launch.  All calls on one GPU share that
 * GPU's array (each clears what it uses first), so calls that overlap on one GPU race. */
ROIFT_API int roift_weights_noalloc_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image,
                                        int image_dtype, void *d_wa, void *d_wb, int weight_dtype, int32_t *h_maxI,
                                        int32_t *h_minI);

/* The same layout for the per-voxel weight forms of the OIFT: base W_p + W_q (d_W, int32 N)
 * or, on Spheric(1.0) only, W_t[p] + W_t[q] with t the arc's axis (d_Wx, d_Wy, d_Wz, int32 N
 * each), times (1 +- per/100) in double by the order of the image values I_p, I_q (unchanged
 * when equal), truncated (gft_ift.cpp:1142-1152, 1244-1254).  Give d_W or all three axes,
 * the others NULL.  Weights must be >= 0; the image may hold any int values.          */
ROIFT_API int roift_weights_w_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                  const int32_t *d_W, const int32_t *d_Wx, const int32_t *d_Wy, const int32_t *d_Wz,
                                  void *d_wa, void *d_wb, int weight_dtype);

/* ---- OIFT under f_max ------------------------------------------------------------ */
/* Drop-in for gft::ift::OIFT_Multi / OIFT_Multi_PerClass (gft_ift.h:108-122,
 * gft_ift.cpp:911-1083).  d_label is in/out exactly as the reference's label
 * scene; the result is bit-identical to the reference's (FIFO tie-breaking
 * included).  d_cost (nullable) receives the f_max cost map
 * C(q) = max(C(pred q), w(pred q -> q)) (INT32_MAX for voxels never reached).  */
ROIFT_API int roift_oift_fmax_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                  const int32_t *d_seeds, int n_seeds, int32_t *d_label, int32_t *d_cost);

/* Drop-in for the per-voxel weight forms gft::ift::OIFT(W, A, scn, per, S, label)
 * (gft_ift.h:125-130, gft_ift.cpp:1086-1167) and OIFT(Wx, Wy, Wz, scn, per, S, label)
 * (gft_ift.h:133-139, gft_ift.cpp:1170-1271): arc weights as roift_weights_w_dev builds
 * them, label > 0 -> +per, label 0 -> -per, the same Prim order and FIFO ties as
 * roift_oift_fmax_dev, so labels and the f_max cost map are the reference's bit for bit.
 * pb->use_per_class must be 0.  (gft::ift::OIFT(A, scn, per, S, label), gft_ift.h:101-106,
 * is OIFT_Multi for labels >= 0: roift_oift_fmax_dev.)                                    */
ROIFT_API int roift_oift_w_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                               const int32_t *d_W, const int32_t *d_Wx, const int32_t *d_Wy, const int32_t *d_Wz,
                               const int32_t *d_seeds, int n_seeds, int32_t *d_label, int32_t *d_cost);

/* Same forest on caller-supplied oriented edge weights (layout as
 * roift_weights_dev writes it, ROIFT_U16 or ROIFT_U32).  Labels for which
 * h_pol_sign[label] > 0 (label in [0, max_label]) use w_obj, the others w_bg;
 * with h_pol_sign == NULL label 0 uses w_bg and every other label w_obj.
 * Used for tie-free (rank) weights, which exceed the image-derived range.     */
ROIFT_API int roift_oift_edges_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_wa, const void *d_wb,
                                   int weight_dtype, const int8_t *h_pol_sign, const int32_t *d_seeds, int n_seeds,
                                   int32_t *d_label, int32_t *d_cost);

/* ---- fuzzy relaxation -------------------------------------------------------------- */
/* Drop-in for gft::ift::ORelax_1_Multi / ORelax_1_Multi_PerClass
 * (gft_ift.h:451-465, gft_ift.cpp:8655-9187): d_label in = OIFT labels,
 * out = argmax labels.  FP32 arithmetic.  d_memb (nullable) receives the final
 * memberships as planar [nclasses][N] floats, classes in increasing label
 * order; h_nclasses (nullable) receives the class count.                      */
ROIFT_API int roift_relax_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                              const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter, float *d_memb,
                              int32_t *h_nclasses);

/* Drop-in for gft::ift::ORelax_1, the binary fallback ORelax_1_Multi takes when S
 * is empty (gft_ift.cpp:8678-8682), and for its per-voxel-weight forms
 * ORelax_1(W, ...) and ORelax_1(Wx, Wy, Wz, ...) (gft_ift.h:443-482,
 * gft_ift.cpp:8498-8653, 9189-9498).  One channel f = label/Imax, hard polarity,
 * thresholded at 0.5: d_label in = labels, out = 0/1.  Weight base: the image
 * (d_W = d_Wx = NULL), int32 d_W (W_p + W_q), or int32 d_Wx/d_Wy/d_Wz (per axis,
 * Spheric(1.0) whatever pb->adj_radius says).  Seeds (n_seeds may be 0) keep their
 * initial value.  d_memb (nullable) receives f (N floats).  roift_relax_dev with
 * n_seeds == 0 runs the image form, as ORelax_1_Multi does.                      */
ROIFT_API int roift_relax_binary_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                     const int32_t *d_W, const int32_t *d_Wx, const int32_t *d_Wy,
                                     const int32_t *d_Wz, const int32_t *d_seeds, int n_seeds, int32_t *d_label,
                                     int niter, float *d_memb);

/* ---- experiment variants (experiments/exp_gradient_weight.cpp, exp_gaussian_rbf_relax.cpp) */
/* OIFT_Multi_GradWeight (exp_gradient_weight.cpp:196-270): the OIFT with arc weight
 * (int)(wf (1 +- per/100)), wf = alpha|I_p-I_q| + (1-alpha)(G_p+G_q)/2 in float,
 * clamped at 0.  d_grad: int32 normalised gradient image (the experiment scales a
 * Sobel magnitude to [0, max I]).  pb->per only (no per-class form).  Exact.     */
ROIFT_API int roift_oift_gradweight_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                        const int32_t *d_grad, float alpha, const int32_t *d_seeds, int n_seeds,
                                        int32_t *d_label, int32_t *d_cost);
/* ORelax_1_Multi_GradWeight (exp_gradient_weight.cpp:272-504): ORelax_1_Multi
 * with the blended base weight, a clamp at 0 after the polarity factor and a
 * float Wmax.  Arguments as roift_relax_dev; leaves d_label unchanged when
 * n_seeds == 0, as the experiment does.                                        */
ROIFT_API int roift_relax_gradweight_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image,
                                         int image_dtype, const int32_t *d_grad, float alpha, const int32_t *d_seeds,
                                         int n_seeds, int32_t *d_label, int niter, float *d_memb);
/* ORelax_1_Multi_GaussRBF (exp_gaussian_rbf_relax.cpp:211-427): w = exp(-d^2 /
 * sigma^2) / Dpq, no polarity.  sigma <= 0 uses estimate_sigma
 * (exp_gaussian_rbf_relax.cpp:194-208, the image's global standard deviation);
 * h_sigma (nullable) receives the sigma used.  Leaves d_label unchanged when
 * n_seeds == 0.                                                                  */
ROIFT_API int roift_relax_rbf_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                  float sigma, const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter,
                                  float *d_memb, float *h_sigma);

/* ---- the whole hot path ------------------------------------------------------------- */
/* OIFT_Multi followed by ORelax_1_Multi on the same inputs, as the reference
 * CLI runs them (oiftrelax.cpp:379-401): weights + OIFT + relax + labels.     */
ROIFT_API int roift_pipeline_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                 const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter,
                                 int32_t *d_cost, float *d_memb);

/* Batched form: OIFT_Multi + ORelax_1_Multi (6-adjacency, polarity `per`) over nvol
 * independent volumes stored back to back in d_image / d_label / d_cost (flat, volume v
 * of h_dims[3v..3v+2] = xs, ys, zs voxels).  d_seeds holds each volume's seeds as
 * volume-local indices, volume after volume, h_nseeds[v] of them.  Every volume's result
 * equals roift_pipeline_dev on that volume alone (bit for bit).  One exact OIFT launch
 * serves the whole batch; the relaxation runs one launch per iteration for all volumes
 * (seed labels {0, L}, intensities < 32767; otherwise volume by volume).  d_cost nullable.
 * use_graph != 0: the whole pipeline (union layout, weights, the cooperative OIFT kernel,
 * gather, relaxation and the status copies) is captured as one HIP graph after the first
 * call with a given key (volume shapes, seed counts, buffer addresses, polarity, iteration
 * count, tuning parameters) and replayed by the calls that repeat that key, with no host
 * synchronisation inside; the checks (seeds, negative images, OIFT errors) are read back
 * after the replay, and weights beyond 16 bits or seed labels other than {0, L} rerun the
 * call eagerly from the saved input labels.  The first call of a key runs eagerly.  No
 * reference counterpart: the reference processes one volume per call
 * (oiftrelax.cpp:379-401).                                                          */
ROIFT_API int roift_batch_pipeline_dev(roift_ctx *ctx, int nvol, const int32_t *h_dims, float adj_radius, float per,
                                       const void *d_image, int image_dtype, const int32_t *d_seeds,
                                       const int32_t *h_nseeds, int32_t *d_label, int niter, int32_t *d_cost,
                                       int use_graph);

/* State of the batched pipeline's HIP graph in this context: graphs captured, replays, replays
 * that reran eagerly (weights beyond 16 bits, seed labels other than {0, L}), and, in err
 * (errlen bytes, nullable), why the last capture failed ("" if none).  No reference
 * counterpart (the reference has no batched form).                                   */
ROIFT_API int roift_batch_graph_info(const roift_ctx *ctx, int *captures, int *replays, int *reruns, char *err,
                                     int errlen);

/* Differentiable two-class relaxation and its adjoint.  No reference counterpart: the
 * learnable weight joins ORelax_1_Multi_GaussRBF's Gaussian (exp_gaussian_rbf_relax.cpp:
 * 343-345) with ORelax_1_Multi's soft polarity (gft_ift.cpp:8800-8812) and a per-voxel
 * scale:  w_pq = scale_q exp(-(I_p-I_q)^2/sigma^2) (1 + alpha sgn(I_p-I_q)(2 f_q - 1)) / Dpq,
 * f = background membership (0 on label == lab_obj, 1 elsewhere at the start), seeds
 * fixed, niter Jacobi iterations.  Loss = binary cross-entropy of o = 1 - f against
 * d_target (N floats in [0,1]), o clamped to [eps, 1-eps].  Returns the loss and
 * dLoss/dscale (d_grad_scale, N floats), dLoss/dsigma, dLoss/dalpha by reverse mode,
 * storing one state every ckpt_every iterations and recomputing each segment.
 * d_memb (nullable) receives the final f.                                          */
ROIFT_API int roift_relax_adjoint_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                      const int32_t *d_seeds, int n_seeds, const int32_t *d_label, int lab_obj,
                                      int niter, int ckpt_every, const float *d_scale, float sigma, float alpha,
                                      const float *d_target, float eps, double *h_loss, float *d_grad_scale,
                                      double *h_grad_sigma, double *h_grad_alpha, float *d_memb);
/* The same in FP64 (sigma, alpha, eps, states, arithmetic, d_grad_scale and d_memb are
 * double; the scale map and the target stay FP32 inputs): the gradient a caller checks
 * against FP64 finite differences (1e-4 relative).                                     */
ROIFT_API int roift_relax_adjoint64_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image,
                                        int image_dtype, const int32_t *d_seeds, int n_seeds, const int32_t *d_label,
                                        int lab_obj, int niter, int ckpt_every, const float *d_scale, double sigma,
                                        double alpha, const float *d_target, double eps, double *h_loss,
                                        double *d_grad_scale, double *h_grad_sigma, double *h_grad_alpha,
                                        double *d_memb);

/* ---- preprocessing ------------------------------------------------------------ */
/* gft::Scene32::GaussianBlur (gft_filtering3.cpp:385-397) applied `passes` times
 * (the CLI blurs twice before the OIFT, oiftrelax.cpp:368-375): 7-tap spherical
 * Gaussian of radius 1 (Kernel3::SphericalGaussian(scn, 1, 10, 1) normalised,
 * gft_filtering3.cpp:72-126), taps outside the volume skipped, FP32 sum in the
 * reference's tap order and gft's ROUND, so the output is bit-exact.  Only
 * pb->xsize/ysize/zsize are read.  d_in and d_out hold N values of image_dtype
 * (I16 or I32; the output has the input's dtype) and may be the same buffer.     */
ROIFT_API int roift_gaussian_blur_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_in, int image_dtype,
                                      void *d_out, int passes);

/* Gradient source of the experiments (exp_gradient_weight.cpp:593-607):
 * gft::Scene32::SobelFilter (gft_filtering3.cpp:509-552) of the image into d_grad
 * (N int32; the reference's linear range [1 + xs + xs*ys, N - 1 - xs - xs*ys) with
 * its wrap-around neighbours, 0 elsewhere), bit-exact; with normalize != 0 then
 * scaled as G = (int)((long long)G * max(image) / max(G)) when both are positive.  */
ROIFT_API int roift_sobel_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                              int32_t *d_grad, int normalize);

/* compute_bfs_edt (exp_geodesic_seeds.cpp:165-215): 6-connected BFS distance from
 * the seeds (plain list) whose label is > 0, into d_dist (N int32; INT_MAX where
 * unreached).  Computed exactly as the L1 distance by separable passes.           */
ROIFT_API int roift_bfs_edt_dev(roift_ctx *ctx, const roift_problem *pb, const int32_t *d_label,
                                const int32_t *d_seeds, int n_seeds, int32_t *d_dist);

/* Background seeds on the volume faces, in the reference's raster order: with
 * geodesic == 0 the CLI's inject_boundary_seeds (oiftrelax.cpp:175-222: voxels whose
 * x, y and z are multiples of stride), with geodesic != 0 inject_boundary_seeds_geodesic
 * (exp_geodesic_seeds.cpp:221-283, distances from d_seeds as roift_bfs_edt_dev).
 * Unlabelled (NIL) candidates get label 0 in d_label and are written to d_new_seeds
 * (append them to the seed list after the caller's seeds); *h_added = their count.
 * ROIFT_EINVAL, with nothing modified, if more than `capacity` would be added.       */
ROIFT_API int roift_boundary_seeds_dev(roift_ctx *ctx, const roift_problem *pb, int32_t *d_label, int stride,
                                       int geodesic, const int32_t *d_seeds, int n_seeds, int32_t *d_new_seeds,
                                       int capacity, int *h_added);

/* The CLI's binary post-process dilation_conditional(fscn, label, radius, percentile)
 * (oiftrelax.cpp:87-169, 403-409), radius = pb->adj_radius: label-0 voxels with a
 * Spheric(radius) neighbour of label 1 form the border; stably sorted by image value,
 * its first (int)(float)((percentile / 100.0) * (count - 1)) + 1 voxels get label 1.
 * *h_set = that number.  An empty border changes nothing (the reference indexes out of
 * range there).                                                                      */
ROIFT_API int roift_dilation_conditional_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image,
                                             int image_dtype, int32_t *d_label, float percentile, int *h_set);

/* ---- volume, seed and polarity files (the CLI's argv contract) ------------------- */
/* Formats: SCN (gft's raw scene, 8/16/32-bit; gft_scene32.cpp:334-419, 602-664) and
 * NIfTI-1 (.nii, .nii.gz, .hdr/.img; uint8, int8, int16, uint16, int32 and float32 as
 * gft::Scene32::ReadNifti1 accepts them, gft_scene32.cpp:525-597; float32 becomes
 * ROUND(v * (10000 / max)) as there).  Host-only entries report errors through
 * roift_io_last_error().                                                            */
enum roift_volume_format { ROIFT_FMT_SCN = 1, ROIFT_FMT_NIFTI = 2 };
typedef struct {
    int xsize, ysize, zsize;
    float dx, dy, dz;          /* voxel spacing */
    int stored_dtype;          /* ROIFT_U8 / I8 / U16 / I16 / I32 / F32, as stored */
    int format;                /* roift_volume_format */
} roift_volume_info;
enum { ROIFT_READ_SHIFT_MIN = 1 };   /* add -min when the minimum is negative (oiftrelax.cpp:263-268) */

ROIFT_API const char *roift_io_last_error(void);
/* Header only. */
ROIFT_API int roift_volume_info_read(const char *path, roift_volume_info *info);
/* gft::Scene32::Read (gft_scene32.cpp:334-419): the N samples as int32. */
ROIFT_API int roift_volume_read(const char *path, int32_t *h_data, long long n);
/* gft::Scene32::Write of a label volume (gft_scene32.cpp:602-664): .scn (8/16/32 bits by
 * the maximum) or int32 NIfTI-1 (.nii / .nii.gz); like_path (nullable) is the input
 * volume, whose NIfTI header fields are kept as gft keeps the cloned header.          */
ROIFT_API int roift_volume_write(const char *path, const int32_t *h_data, const roift_volume_info *info,
                                 const char *like_path);
/* The CLI's seeds file (oiftrelax.cpp:323-351): "<n>" then n lines "x y z id label";
 * voxels outside the volume are skipped, negative labels become 0 (*h_clamped = 1).
 * h_seeds / h_labels receive up to capacity entries in file order; *h_count = all
 * (with both null, it only counts).                                                 */
ROIFT_API int roift_seeds_read(const char *path, int xsize, int ysize, int zsize, int32_t *h_seeds,
                               int32_t *h_labels, int capacity, int *h_count, int *h_clamped);
/* The CLI's per-class polarity file (oiftrelax.cpp:271-310): "<n>" then n values in
 * [-1, 1]; h_per_class receives them times 100.0f.                                   */
ROIFT_API int roift_polarity_read(const char *path, float *h_per_class, int capacity, int *h_count);
/* Reads a volume straight into device memory (d_dst: N values of dst_dtype, I16 or I32):
 * the raw samples go through two pinned host buffers, double-buffered against the file
 * reads, and are converted on the device.  With ROIFT_READ_SHIFT_MIN a negative minimum
 * is added back (*h_shift = the amount).  ROIFT_EINVAL when the values do not fit I16.  */
ROIFT_API int roift_volume_read_dev(roift_ctx *ctx, const char *path, void *d_dst, int dst_dtype, int flags,
                                    int *h_shift, roift_volume_info *info);

/* The reference CLI (oiftrelax.cpp:227-432) with every stage on the GPU: argv as
 * oiftrelax's (argv[0] = program name, then <volume> <file_seeds> <pol> <niter>
 * <percentile> <output_file> [boundary_stride] [pol_file]); prints its progress and
 * stage timings to stdout.  Returns 0, or 1 after printing the failing stage to stderr.
 * roift_amd/roift_oiftrelax is this function as an executable.                        */
ROIFT_API int roift_cli_main(int argc, char **argv);

/* Host-pointer forms (copy in, run, copy out).  h_image is int32 as gft's Scene32. */
ROIFT_API int roift_oift_fmax(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_image,
                              const int32_t *h_seeds, int n_seeds, int32_t *h_label, int32_t *h_cost);
ROIFT_API int roift_oift_w(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_image, const int32_t *h_W,
                           const int32_t *h_Wx, const int32_t *h_Wy, const int32_t *h_Wz, const int32_t *h_seeds,
                           int n_seeds, int32_t *h_label, int32_t *h_cost);
ROIFT_API int roift_relax(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_image,
                          const int32_t *h_seeds, int n_seeds, int32_t *h_label, int niter, float *h_memb);
ROIFT_API int roift_relax_binary(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_image, const int32_t *h_W,
                                 const int32_t *h_Wx, const int32_t *h_Wy, const int32_t *h_Wz,
                                 const int32_t *h_seeds, int n_seeds, int32_t *h_label, int niter, float *h_memb);
ROIFT_API int roift_gaussian_blur(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_in, int32_t *h_out,
                                  int passes);

/* ---- legacy reference C-ABI ---------------------------------------------------------- */
/* Same symbols and signatures as the reference's two GPU entry points.  Both compute
 * OIFT_Multi_PerClass with the given per-class polarities exactly (6-adjacency), under the
 * reference's input contract: every voxel that is not a listed seed starts unlabelled,
 * whatever h_labels holds; a seed's label is h_labels[seed] with negative labels read as 0;
 * seeds outside [0, n) are skipped and a repeated seed counts once (its first position is
 * its FIFO position); on return every voxel holds its label, 0 where no seed reached it.
 * Return: the number of level steps (the reference returns its wave count), or -1 on
 * error (message on stderr).  Per-class tables with more than two polarities (other than
 * +P / -P) are outside this library and return -1.
 *
 * oift_fmax_gpu_run: gpu/oift_fmax_gpu.cu:123-147, declared by its caller at
 * gpu/oiftrelax_fmax_gpu.cpp:21-25.  With no valid seed every label becomes 0 (the
 * reference's untouched (COST_INF, 0) words) and it returns 0.                          */
ROIFT_API int oift_fmax_gpu_run(const int *h_image, int *h_labels, int n, int xsize, int ysize, int zsize,
                                const float *h_per_class, int max_label, const int *h_seeds, int n_seeds);
/* oift_gpu_run: gpu/oift_gpu.cu:149-158, declared by its caller at gpu/oiftrelax_gpu.cpp:22-28
 * and called at :184-189.  cost_mode 0 = f_max (the path this library implements); any
 * other mode (1 = additive) returns -1.  Seed labels above max_label are read as 0
 * (oift_gpu.cu:202); with no valid seed h_labels is left as it is and it returns 0
 * (oift_gpu.cu:207-210).                                                                */
ROIFT_API int oift_gpu_run(int cost_mode, const int *h_image, int *h_labels, int n, int xsize, int ysize,
                           int zsize, const float *h_per_class, int max_label, const int *h_seeds, int n_seeds);

#ifdef __cplusplus
}
#endif

#endif /* ROIFT_H */
