// ctx.hpp -- per-device context: stream, cached device buffers, pinned counters, timings.
#pragma once

#include <map>
#include <string>

#include "common.hpp"

struct roift_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    std::string last_error;
    std::map<std::string, std::pair<void *, size_t>> bufs;
    uint32_t *h_ctr = nullptr;      // pinned host mirror of the device counters
    double timings[6] = {0, 0, 0, 0, 0, 0};
    double prof[180] = {};          // OIFT phase ms [0..8], counts [9..17], ms by log2(entries) [18..161],
                                    // flood items [162], touched slots [163], batches sum [164] / max per workgroup [165],
                                    // sum over flood rounds of the busiest workgroup's batches [166],
                                    // flood batches by ceil(log2(items)) [167..176], items of the rounds' busiest
                                    // workgroups [177], single-wave flood hops [178] and their ms [179]
    int tune_blocks = 0;            // OIFT persistent grid size (0: sized from the arc count)
    int tune_budget = 0;            // OIFT LDS batches per workgroup per flood round (0: 1024)
    int tune_relax_depth = 0;       // reserved (ROIFT_PARAM_RELAX_DEPTH is accepted and ignored)
    int tune_flood_mode = 0;        // OIFT flood: 0 = single voxels, 1 = 8x8x8 bricks (6-adjacency)
    int tune_relax_tile = 0;        // z-marching relaxation tile: rows | planes << 8 | xcd << 16 (0: default)
    int tune_blur_planes = 0;       // Gaussian blur: planes per workgroup (0: automatic)
    int tune_oift_mode = 0;         // OIFT over explicit weights: 0 level-synchronous, 1 local steps (tie-free maps)
    int tune_local_cand = 0;        // local steps: candidates per round (0 = 256)
    int tune_local_steps = 0;       // local steps: steps a check takes before it gives up (0 = 256)
    int tune_oift_profile = 0;      // 1: the OIFT's flood critical-path / batch / hop counters
    int oift_local_fallbacks = 0;   // local-step runs that met a tie and reran level-synchronously
    hipEvent_t ev[16] = {};
    // the batched pipeline as one HIP graph (scatter, weights, OIFT, gather, relaxation, status
    // copies), replayed while its key (shapes, pointers, parameters) and every context buffer it
    // captured are unchanged (buf_gen counts the reallocations)
    hipGraphExec_t batch_graph = nullptr;
    std::string batch_graph_key;
    unsigned long long batch_graph_gen = 0;
    unsigned long long buf_gen = 0;
    int *h_bstat = nullptr;         // pinned: the graph's status words (batch info, weight flags)
    size_t h_bstat_n = 0;
    int batch_captures = 0, batch_replays = 0, batch_reruns = 0;   // graph builds, replays, eager reruns
    std::string batch_graph_error;  // why the last capture failed (the call then ran eagerly)
    std::string batch_graph_fail_key;
    void *pinned[2] = {nullptr, nullptr};   // pinned host staging for volume reads (double-buffered)
    size_t pinned_bytes = 0;
    hipEvent_t io_ev[2] = {};
    // device-event times of the last call's main kernels, ms: weight builder, OIFT
    // persistent kernel, all relaxation iterations; [3] = relaxation launches; [4] = Gaussian blur;
    // [5] = the last CLI stage call (Sobel, distance, boundary seeds or dilation), all its kernels
    double ktimes[6] = {0, 0, 0, 0, 0, 0};
    // this device's copy of the allocation-free weight builder's static scratch array,
    // resolved at creation (nullptr if hipGetSymbolAddress failed; that entry then errors)
    int *noalloc_scratch = nullptr;
    std::string noalloc_error;
};

namespace roift {

// Device buffer owned by the context, grown on demand and reused across calls.
template <typename T>
T *buf(roift_ctx &c, const std::string &name, size_t count)
{
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 256;
    auto it = c.bufs.find(name);
    if (it != c.bufs.end() && it->second.second >= bytes) return static_cast<T *>(it->second.first);
    if (it != c.bufs.end()) {
        ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
        ROIFT_HIP_CHECK(hipFree(it->second.first));
        c.bufs.erase(it);
        c.buf_gen++;   // a captured graph that held the old address is stale
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) throw Error{ROIFT_ENOMEM, "hipMalloc failed for " + name};
    c.bufs[name] = {p, bytes};
    return static_cast<T *>(p);
}

// Device events on the context's stream (slots 0-7: stages, 8-15: kernels).
inline void ev_record(roift_ctx &c, int i) { ROIFT_HIP_CHECK(hipEventRecord(c.ev[i], c.stream)); }

inline double ev_ms(roift_ctx &c, int a, int b)
{
    float ms = 0.f;
    ROIFT_HIP_CHECK(hipEventSynchronize(c.ev[b]));
    ROIFT_HIP_CHECK(hipEventElapsedTime(&ms, c.ev[a], c.ev[b]));
    return ms;
}

inline unsigned grid_for(long long n, int block = 256, long long cap = 8192)
{
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

template <typename T>
void fill(roift_ctx &c, T *p, long long n, T v)
{
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(256), 0, c.stream, p, n, v);
    ROIFT_HIP_CHECK(hipGetLastError());
}

// Stage results of the oriented weight builder.
struct WeightInfo {
    int32_t maxI = 0, minI = 0;
};

// Entry points implemented in the .hip translation units.
// Optional base of the builder's arc weight in place of |I_p - I_q|: the gradient-weight
// model (OIFT_Multi_GradWeight: int32 G, blend alpha), a per-voxel weight W (OIFT(W, ...):
// W_p + W_q) or per-axis weights Wx, Wy, Wz (OIFT(Wx, Wy, Wz, ...), Spheric(1.0) only).
struct GradModel {
    const int32_t *G = nullptr;
    float alpha = 0.5f;
    const int32_t *W = nullptr;
    const int32_t *Wx = nullptr, *Wy = nullptr, *Wz = nullptr;
};

// d_scratch (nullable): the allocation-free form's scratch (noalloc_scratch); otherwise the
// extrema partials come from context buffers.
WeightInfo build_weights(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, float per,
                         void *d_wa, void *d_wb, int weight_dtype, const GradModel *gm = nullptr,
                         int *d_scratch = nullptr);
// The builder's launches alone (no host read-back: graph capture): ext (4 ints) receives the
// image extrema and the overflow flag, part two ints per workgroup (weights_blocks(d) of them).
long long weights_blocks(const Dims &d);
void enqueue_weights(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, float per,
                     void *d_wa, void *d_wb, int weight_dtype, const GradModel *gm, int *ext, int *part, bool timed);
// Weight-builder workgroups the allocation-free entry's static scratch holds (8 MiB per device).
constexpr long long NOALLOC_MAX_BLOCKS = 1 << 20;
// This device's copy of that scratch array, resolved once per device with hipGetSymbolAddress
// (call with `device` current).
int *noalloc_scratch(int device);
WeightInfo image_extrema(roift_ctx &c, const Dims &d, const void *d_image, int image_dtype);
// Global standard deviation of the image (estimate_sigma, exp_gaussian_rbf_relax.cpp:194-208).
float image_sigma(roift_ctx &c, const Dims &d, const void *d_image, int image_dtype);

// Class of each label for the OIFT: 1 = w_obj, 0 = w_bg (transposed).  Labels
// outside [0, max_label] use default_cls.
struct LabelClasses {
    std::vector<int8_t> cls;  // index = label
    int default_cls = 1;
};

int run_oift(roift_ctx &c, const Dims &d, int nadj, const void *d_wa, const void *d_wb, int weight_dtype,
             const LabelClasses &lc, const int32_t *d_seeds, int n_seeds, int32_t *d_label, int32_t *d_cost);
// The same OIFT in three stages for a HIP graph: OIFT_PREPARE allocates every buffer and uploads
// the class table (before a capture), OIFT_ENQUEUE issues the launches and the counter copy with
// no host synchronisation (seed errors stop the kernel on the device), OIFT_FINISH (after the
// graph has run) reports errors and fills the profile.  n_seeds > 0.
enum : int { OIFT_PREPARE = 0, OIFT_ENQUEUE = 1, OIFT_FINISH = 2 };
int run_oift_staged(roift_ctx &c, int stage, const Dims &d, int nadj, const void *d_wa, const void *d_wb,
                    int weight_dtype, const LabelClasses &lc, const int32_t *d_seeds, int n_seeds, int32_t *d_label,
                    int32_t *d_cost);
// OIFT_Multi's forest from the image: the arc weights, their packed slots and the OIFT state
// in one fused pass where the weights fit 16 bits (planar weights otherwise); wi receives the
// image extrema.
int run_oift_image(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, float per,
                   const LabelClasses &lc, const int32_t *d_seeds, int n_seeds, int32_t *d_label, int32_t *d_cost,
                   WeightInfo &wi);

struct RelaxSpec {
    int use_per_class = 0;
    float per = 0.f;
    std::vector<float> per_class;  // indexed by label, length max_label+1
    int max_label = -1;
};

// Relaxation weight models: ORelax_1_Multi[_PerClass], ORelax_1_Multi_GradWeight, ORelax_1_Multi_GaussRBF.
enum : int { RELAX_INTENSITY = 0, RELAX_GRAD = 1, RELAX_RBF = 2 };
struct RelaxModel {
    int model = RELAX_INTENSITY;
    const int32_t *G = nullptr;   // RELAX_GRAD: normalised gradient image (int32, device)
    float alpha = 0.5f;           // RELAX_GRAD: blend
    float sigma = 1.0f;           // RELAX_RBF
};

int run_relax(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, const RelaxSpec &rs,
              const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter, float *d_memb, int32_t *nclasses_out,
              const WeightInfo &wi, const RelaxModel &rm = RelaxModel());

// One volume of a batch stored back to back: voxel offset, size and image extrema.
struct BatchVolume {
    long long off;
    int xs, ys, zs;
    int32_t maxI, minI;
};

// gft::Scene32::GaussianBlur applied `passes` times (in place when d_in == d_out).
int run_gaussian_blur(roift_ctx &c, const Dims &d, const void *d_in, int dtype, void *d_out, int passes);

// Volume, seed and polarity files (volio.hip).  Host-only entries report errors through
// io_last_error(); the device read also sets the context's last error.
const char *io_last_error();
int volume_info(const char *path, roift_volume_info *info);
int volume_read(const char *path, int32_t *out, long long n);
int volume_write(const char *path, const int32_t *data, const roift_volume_info *info, const char *like);
int seeds_read(const char *path, int xs, int ys, int zs, int32_t *seeds, int32_t *labels, int cap, int *count,
               int *clamped);
int polarity_read(const char *path, float *per_class, int cap, int *count);
int volume_read_dev(roift_ctx &c, const char *path, void *d_dst, int dst_dtype, int flags, int *h_shift,
                    roift_volume_info *info);

// The CLI's stages around the hot path (stages.hip).
int run_sobel(roift_ctx &c, const Dims &d, const void *d_image, int dtype, int32_t *d_grad, bool normalize);
int run_bfs_edt(roift_ctx &c, const Dims &d, const int32_t *d_label, const int32_t *d_seeds, int n_seeds,
                int32_t *d_dist);
int run_boundary_seeds(roift_ctx &c, const Dims &d, int32_t *d_label, int stride, bool geodesic,
                       const int32_t *d_seeds, int n_seeds, int32_t *d_out, int capacity, int *h_added);
int run_dilation_conditional(roift_ctx &c, const Dims &d, float radius, const void *d_image, int dtype,
                             int32_t *d_label, float percentile, int *h_set);

// Two-class 6-adjacency ORelax_1_Multi over a batch in one launch per iteration: d_lab_obj (device)
// is the object label (classes {0, lab_obj}), d_ext (device) each volume's image max and min;
// d_seeds are flat indices.  mode: RELAX_BATCH_EAGER uploads the tile and volume tables and runs;
// RELAX_BATCH_CAPTURE only launches (a HIP graph records them); RELAX_BATCH_TABLES only allocates
// and uploads the tables (they depend on the shapes only: before a capture and before each replay).
enum : int { RELAX_BATCH_EAGER = 0, RELAX_BATCH_CAPTURE = 1, RELAX_BATCH_TABLES = 2 };
int run_relax_batch(roift_ctx &c, const std::vector<BatchVolume> &vols, const void *d_image, int image_dtype,
                    float per, const int *d_lab_obj, const int *d_ext, const int32_t *d_seeds, int n_seeds,
                    int32_t *d_label, int niter, int mode);

// OIFT_Multi + ORelax_1_Multi (6-adjacency, one polarity) over nvol independent volumes
// stored back to back: h_dims = xs, ys, zs per volume; d_seeds = each volume's seeds as
// volume-local indices, volume after volume (h_nseeds[v] of them); d_label / d_cost flat.
int run_batch(roift_ctx &c, int nvol, const int32_t *h_dims, float adj_r, float per, const void *d_image,
              int image_dtype, const int32_t *d_seeds, const int32_t *h_nseeds, int32_t *d_label, int niter,
              int32_t *d_cost, bool use_graph);

// Differentiable two-class relaxation and its checkpointed adjoint (learnable per-voxel
// scale, RBF sigma and polarity alpha; BCE loss against d_target).  fp64: states,
// gradients and memberships in double (d_gscale / d_memb are double *), else float.
int run_relax_diff(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype,
                   const int32_t *d_seeds, int n_seeds, const int32_t *d_label, int lab_obj, int niter, int ckpt,
                   const float *d_scale, double sigma, double alpha, const float *d_target, double eps,
                   double *h_loss, void *d_gscale, double *h_gsig, double *h_galp, void *d_memb, bool fp64);

// ORelax_1 (binary fallback): single channel f = label/Imax, hard polarity.  Weight source:
// the image (W == Wx == nullptr), a per-voxel weight W, or per-axis Wx/Wy/Wz (6-adj only).
int run_relax_binary(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, float per,
                     const int32_t *W, const int32_t *Wx, const int32_t *Wy, const int32_t *Wz,
                     const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter, float *d_memb);

}  // namespace roift
