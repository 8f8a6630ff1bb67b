// capi.hip -- the extern "C" boundary of libroift.so (declared in include/roift.h).
#include <algorithm>
#include <cstring>
#include <cmath>
#include <cstdio>
#include <set>
#include <vector>

#include "ctx.hpp"

using namespace roift;

namespace {

// Spheric(r) neighbourhood size including the centre (gft_adjrel3.cpp:47-85).
int spheric_n(float r)
{
    const int r0 = (int)r, r2 = (int)(r * r);
    int n = 0;
    for (int z = -r0; z <= r0; z++)
        for (int y = -r0; y <= r0; y++)
            for (int x = -r0; x <= r0; x++)
                if (x * x + y * y + z * z <= r2) n++;
    return n;
}

Dims dims_of(const roift_problem *pb)
{
    ROIFT_REQUIRE(pb != nullptr, ROIFT_EINVAL, "null problem");
    ROIFT_REQUIRE(pb->xsize > 0 && pb->ysize > 0 && pb->zsize > 0, ROIFT_EINVAL, "non-positive volume size");
    Dims d{pb->xsize, pb->ysize, pb->zsize, (long long)pb->xsize * pb->ysize * pb->zsize};
    ROIFT_REQUIRE(d.n < (1ll << 31), ROIFT_EINVAL, "volume larger than 2^31-1 voxels");
    ROIFT_REQUIRE(std::isfinite(pb->dx) && std::isfinite(pb->dy) && std::isfinite(pb->dz), ROIFT_EINVAL,
                  "non-finite voxel spacing");
    d.sx = pb->dx != 0.0f ? pb->dx : 1.0f;
    d.sy = pb->dy != 0.0f ? pb->dy : 1.0f;
    d.sz = pb->dz != 0.0f ? pb->dz : 1.0f;
    return d;
}

int nadj_of(const roift_problem *pb)
{
    const int n = spheric_n(pb->adj_radius);
    ROIFT_REQUIRE(n == 7 || n == 27, ROIFT_EUNSUPPORTED, "only Spheric(1.0) (6-adj) and Spheric(1.75) (26-adj)");
    return n;
}

// Split the labels into the two weight classes of the edge arrays built with
// +P: class 1 uses w_P (arcs as stored), class 0 uses w_{-P} (transposed).
struct Polarity {
    float P = 0.f;
    LabelClasses lc;
};

Polarity polarity_of(const roift_problem *pb)
{
    Polarity po;
    if (!pb->use_per_class) {
        ROIFT_REQUIRE(std::fabs(pb->per) <= 100.0f, ROIFT_EINVAL, "|per| > 100");
        po.P = pb->per;
        po.lc.cls = {0};     // label 0 -> -per (gft_ift.cpp:968)
        po.lc.default_cls = 1;
        return po;
    }
    ROIFT_REQUIRE(pb->h_per_class != nullptr && pb->max_label >= 0, ROIFT_EINVAL, "per_class missing");
    std::set<float> vals;
    for (int c = 0; c <= pb->max_label; c++) {
        ROIFT_REQUIRE(std::fabs(pb->h_per_class[c]) <= 100.0f, ROIFT_EINVAL, "|per_class| > 100");
        vals.insert(pb->h_per_class[c]);
    }
    float P = *vals.rbegin();
    if (vals.size() == 2) {
        ROIFT_REQUIRE(*vals.begin() == -P, ROIFT_EUNSUPPORTED,
                      "per_class with two values that are not +-P is not implemented");
    } else {
        ROIFT_REQUIRE(vals.size() == 1, ROIFT_EUNSUPPORTED, "per_class with more than two polarities is not implemented");
        if (P < 0) P = -P;   // a single value -P: every label uses the transposed arcs of +P
    }
    po.P = P;
    po.lc.cls.resize(pb->max_label + 1);
    for (int c = 0; c <= pb->max_label; c++) po.lc.cls[c] = (pb->h_per_class[c] == P) ? 1 : 0;
    // labels outside [0, max_label] get polarity 0 (gft_ift.cpp:1058-1062)
    po.lc.default_cls = (P == 0.0f) ? 1 : -1;
    return po;
}

RelaxSpec relax_spec_of(const roift_problem *pb)
{
    RelaxSpec rs;
    rs.use_per_class = pb->use_per_class;
    rs.per = pb->per;
    rs.max_label = pb->max_label;
    if (pb->use_per_class) {
        ROIFT_REQUIRE(pb->h_per_class != nullptr && pb->max_label >= 0, ROIFT_EINVAL, "per_class missing");
        rs.per_class.assign(pb->h_per_class, pb->h_per_class + pb->max_label + 1);
    }
    return rs;
}

template <typename F>
int guarded(roift_ctx *ctx, F &&f)
{
    if (!ctx) return ROIFT_EINVAL;
    try {
        ROIFT_HIP_CHECK(hipSetDevice(ctx->device));
        ctx->last_error.clear();
        return f();
    } catch (const Error &e) {
        ctx->last_error = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        return ROIFT_ENOMEM;
    }
}

struct Weights {
    void *wa, *wb;
    int dtype;
    WeightInfo info;
};

Weights weights_into_ctx(roift_ctx &c, const Dims &d, int nadj, const void *img, int idt, float P,
                         const GradModel *gm = nullptr)
{
    const long long slots = (long long)((nadj - 1) / 2) * d.n;
    Weights w{};
    w.dtype = ROIFT_U16;
    w.wa = buf<uint16_t>(c, "pipe.wa", slots);
    w.wb = buf<uint16_t>(c, "pipe.wb", slots);
    try {
        w.info = build_weights(c, d, nadj, img, idt, P, w.wa, w.wb, ROIFT_U16, gm);
    } catch (const Error &e) {
        if (e.code != ROIFT_EUNSUPPORTED) throw;
        w.dtype = ROIFT_U32;
        w.wa = buf<uint32_t>(c, "pipe.wa32", slots);
        w.wb = buf<uint32_t>(c, "pipe.wb32", slots);
        w.info = build_weights(c, d, nadj, img, idt, P, w.wa, w.wb, ROIFT_U32, gm);
    }
    // the image form's weights |I_p - I_q| stay within the reference's queue only for I >= 0
    // (oiftrelax.cpp:263-268 shifts it); the W forms' weights do not depend on its sign
    const bool w_base = gm && !gm->G;
    ROIFT_REQUIRE(w_base || w.info.minI >= 0, ROIFT_EINVAL, "image has negative values (shift it to >= 0 first)");
    return w;
}

// The per-voxel weight forms OIFT(W, ...) / OIFT(Wx, Wy, Wz, ...): W >= 0 (the reference's
// bucket queue has Wmax + 2 buckets, Wmax = 2 max W (1 + |per|/100), gft_ift.cpp:1101-1103,
// 1201-1206), and W_p + W_q (1 + |per|/100) within int32, as the reference's int arithmetic.
GradModel w_model(roift_ctx &c, const Dims &d, int nadj, const int32_t *W, const int32_t *Wx, const int32_t *Wy,
                  const int32_t *Wz, float per)
{
    GradModel gm;
    const bool axes = Wx || Wy || Wz;
    ROIFT_REQUIRE(axes ? (Wx && Wy && Wz && !W) : (W != nullptr), ROIFT_EINVAL,
                  "give either W or all of Wx, Wy, Wz");
    ROIFT_REQUIRE(!axes || nadj == 7, ROIFT_EINVAL, "the per-axis form is defined on Spheric(1.0) only");
    long long wmax = 0;
    for (const int32_t *w : {W, Wx, Wy, Wz}) {
        if (!w) continue;
        const WeightInfo e = image_extrema(c, d, w, ROIFT_I32);
        ROIFT_REQUIRE(e.minI >= 0, ROIFT_EINVAL, "per-voxel weights must be >= 0");
        wmax = std::max<long long>(wmax, e.maxI);
    }
    ROIFT_REQUIRE((double)(2 * wmax) * (1.0 + std::fabs((double)per) / 100.0) < 2147483647.0, ROIFT_EUNSUPPORTED,
                  "per-voxel weights too large for int32 arc weights");
    gm.W = W;
    gm.Wx = Wx;
    gm.Wy = Wy;
    gm.Wz = Wz;
    return gm;
}

}  // namespace

extern "C" {

ROIFT_API int roift_create(roift_ctx **out, int device)
{
    if (!out) return ROIFT_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return ROIFT_EHIP;
    auto *c = new roift_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&c->h_ctr, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return ROIFT_EHIP;
    }
    c->owns_stream = true;
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            roift_destroy(c);
            return ROIFT_EHIP;
        }
    // once per device, with the device current: the address of the allocation-free entry's
    // scratch array on this device (checked; the entry reports a failure here)
    try {
        c->noalloc_scratch = noalloc_scratch(device);
    } catch (const Error &e) {
        c->noalloc_error = e.msg;
    }
    *out = c;
    return ROIFT_OK;
}

ROIFT_API void roift_destroy(roift_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->bufs) (void)hipFree(kv.second.first);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->h_ctr) (void)hipHostFree(c->h_ctr);
    if (c->batch_graph) (void)hipGraphExecDestroy(c->batch_graph);
    if (c->h_bstat) (void)hipHostFree(c->h_bstat);
    for (auto &p : c->pinned)
        if (p) (void)hipHostFree(p);
    for (auto &e : c->io_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->owns_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

ROIFT_API const char *roift_last_error(const roift_ctx *c) { return c ? c->last_error.c_str() : "null context"; }

ROIFT_API int roift_set_stream(roift_ctx *c, void *stream)
{
    return guarded(c, [&]() -> int {
        if (c->owns_stream && c->stream) ROIFT_HIP_CHECK(hipStreamDestroy(c->stream));
        c->stream = (hipStream_t)stream;
        c->owns_stream = false;
        return ROIFT_OK;
    });
}

ROIFT_API int roift_last_timings(const roift_ctx *c, double *out6)
{
    if (!c || !out6) return ROIFT_EINVAL;
    for (int i = 0; i < 6; i++) out6[i] = c->timings[i];
    return ROIFT_OK;
}

ROIFT_API int roift_kernel_timings(const roift_ctx *c, double *out, int n)
{
    if (!c || !out || n < 4) return ROIFT_EINVAL;
    for (int i = 0; i < 4; i++) out[i] = c->ktimes[i];
    if (n >= 5) out[4] = c->ktimes[4];
    if (n >= 6) out[5] = c->ktimes[5];
    return ROIFT_OK;
}

ROIFT_API int roift_set_param(roift_ctx *c, int param, int value)
{
    if (!c || value < 0) return ROIFT_EINVAL;
    switch (param) {
    case ROIFT_PARAM_OIFT_BLOCKS: c->tune_blocks = value; return ROIFT_OK;
    case ROIFT_PARAM_FLOOD_BUDGET: c->tune_budget = value; return ROIFT_OK;
    case ROIFT_PARAM_RELAX_DEPTH: c->tune_relax_depth = value; return ROIFT_OK;
    case ROIFT_PARAM_FLOOD_MODE: c->tune_flood_mode = value; return ROIFT_OK;
    case ROIFT_PARAM_RELAX_TILE: c->tune_relax_tile = value; return ROIFT_OK;
    case ROIFT_PARAM_OIFT_MODE:
        if (value > 1) return ROIFT_EINVAL;
        c->tune_oift_mode = value;
        return ROIFT_OK;
    case ROIFT_PARAM_BLUR_PLANES: c->tune_blur_planes = value; return ROIFT_OK;
    case ROIFT_PARAM_LOCAL_CANDIDATES: c->tune_local_cand = value; return ROIFT_OK;
    case ROIFT_PARAM_LOCAL_CHECK_STEPS: c->tune_local_steps = value; return ROIFT_OK;
    case ROIFT_PARAM_OIFT_PROFILE: c->tune_oift_profile = value; return ROIFT_OK;
    default: return ROIFT_EINVAL;
    }
}

ROIFT_API int roift_oift_profile(const roift_ctx *c, double *out, int n)
{
    if (!c || !out || n < 18) return ROIFT_EINVAL;
    for (int i = 0; i < n && i < 166; i++) out[i] = c->prof[i];
    if (n >= 167) out[166] = c->oift_local_fallbacks;
    if (n >= 168) out[167] = c->prof[166];
    for (int i = 168; i < n && i < 181; i++) out[i] = c->prof[i - 1];
    return ROIFT_OK;
}

ROIFT_API int roift_weights_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                void *d_wa, void *d_wb, int weight_dtype, int32_t *h_maxI, int32_t *h_minI)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        ROIFT_REQUIRE(!pb->use_per_class, ROIFT_EINVAL, "the weight builder takes a single polarity (per)");
        const WeightInfo wi = build_weights(*ctx, d, nadj, d_image, image_dtype, pb->per, d_wa, d_wb, weight_dtype);
        if (h_maxI) *h_maxI = wi.maxI;
        if (h_minI) *h_minI = wi.minI;
        return ROIFT_OK;
    });
}

ROIFT_API int roift_weights_noalloc_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image,
                                        int image_dtype, void *d_wa, void *d_wb, int weight_dtype, int32_t *h_maxI,
                                        int32_t *h_minI)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        ROIFT_REQUIRE(!pb->use_per_class, ROIFT_EINVAL, "the weight builder takes a single polarity (per)");
        ROIFT_REQUIRE(ctx->noalloc_scratch != nullptr, ROIFT_EHIP,
                      "no device address for the scratch array: " + ctx->noalloc_error);
        // the same kernels as roift_weights_dev; the extrema partials and the overflow flag live
        // in the static scratch array instead of context buffers, so nothing is allocated
        const WeightInfo wi = build_weights(*ctx, d, nadj, d_image, image_dtype, pb->per, d_wa, d_wb, weight_dtype,
                                            nullptr, ctx->noalloc_scratch);
        if (h_maxI) *h_maxI = wi.maxI;
        if (h_minI) *h_minI = wi.minI;
        return ROIFT_OK;
    });
}

ROIFT_API int roift_oift_fmax_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                  const int32_t *d_seeds, int n_seeds, int32_t *d_label, int32_t *d_cost)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        const Polarity po = polarity_of(pb);
        ev_record(*ctx, 0);
        WeightInfo wi;
        const int lv = run_oift_image(*ctx, d, nadj, d_image, image_dtype, po.P, po.lc, d_seeds, n_seeds, d_label,
                                      d_cost, wi);
        ev_record(*ctx, 2);
        // stage split: the fused weight / state set-up kernel's time, then the rest of the OIFT
        ctx->timings[0] = ctx->ktimes[0];
        ctx->timings[1] = ev_ms(*ctx, 0, 2) - ctx->timings[0];
        return lv >= 0 ? ROIFT_OK : lv;
    });
}

ROIFT_API int roift_weights_w_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                  const int32_t *d_W, const int32_t *d_Wx, const int32_t *d_Wy, const int32_t *d_Wz,
                                  void *d_wa, void *d_wb, int weight_dtype)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        ROIFT_REQUIRE(!pb->use_per_class, ROIFT_EINVAL, "the weight builder takes a single polarity (per)");
        ROIFT_REQUIRE(std::fabs(pb->per) <= 100.0f, ROIFT_EINVAL, "|per| > 100");
        const GradModel gm = w_model(*ctx, d, nadj, d_W, d_Wx, d_Wy, d_Wz, pb->per);
        build_weights(*ctx, d, nadj, d_image, image_dtype, pb->per, d_wa, d_wb, weight_dtype, &gm);
        return ROIFT_OK;
    });
}

ROIFT_API int roift_oift_w_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                               const int32_t *d_W, const int32_t *d_Wx, const int32_t *d_Wy, const int32_t *d_Wz,
                               const int32_t *d_seeds, int n_seeds, int32_t *d_label, int32_t *d_cost)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        ROIFT_REQUIRE(!pb->use_per_class, ROIFT_EINVAL, "OIFT(W, ...) takes a single polarity (per)");
        const Polarity po = polarity_of(pb);   // label > 0 -> +per, label 0 -> -per (gft_ift.cpp:1140-1143)
        const GradModel gm = w_model(*ctx, d, nadj, d_W, d_Wx, d_Wy, d_Wz, po.P);
        ev_record(*ctx, 0);
        const Weights w = weights_into_ctx(*ctx, d, nadj, d_image, image_dtype, po.P, &gm);
        ev_record(*ctx, 1);
        const int lv = run_oift(*ctx, d, nadj, w.wa, w.wb, w.dtype, po.lc, d_seeds, n_seeds, d_label, d_cost);
        ev_record(*ctx, 2);
        ctx->timings[0] = ev_ms(*ctx, 0, 1);
        ctx->timings[1] = ev_ms(*ctx, 1, 2);
        return lv >= 0 ? ROIFT_OK : lv;
    });
}

ROIFT_API int roift_oift_edges_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_wa, const void *d_wb,
                                   int weight_dtype, const int8_t *h_pol_sign, const int32_t *d_seeds, int n_seeds,
                                   int32_t *d_label, int32_t *d_cost)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        LabelClasses lc;
        if (h_pol_sign) {
            ROIFT_REQUIRE(pb->max_label >= 0, ROIFT_EINVAL, "pol_sign needs max_label");
            lc.cls.resize(pb->max_label + 1);
            for (int c = 0; c <= pb->max_label; c++) lc.cls[c] = h_pol_sign[c] > 0 ? 1 : 0;
            lc.default_cls = 1;
        } else {
            lc.cls = {0};
            lc.default_cls = 1;
        }
        ev_record(*ctx, 1);
        const int lv = run_oift(*ctx, d, nadj, d_wa, d_wb, weight_dtype, lc, d_seeds, n_seeds, d_label, d_cost);
        ev_record(*ctx, 2);
        ctx->timings[1] = ev_ms(*ctx, 1, 2);
        return lv >= 0 ? ROIFT_OK : lv;
    });
}

ROIFT_API int roift_relax_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                              const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter, float *d_memb,
                              int32_t *h_nclasses)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        const RelaxSpec rs = relax_spec_of(pb);
        ev_record(*ctx, 2);
        if (n_seeds == 0) {   // ORelax_1_Multi with an empty S runs ORelax_1 (gft_ift.cpp:8678-8682)
            ROIFT_REQUIRE(!pb->use_per_class, ROIFT_EINVAL, "the per-class relaxation needs seeds");
            run_relax_binary(*ctx, d, nadj, d_image, image_dtype, pb->per, nullptr, nullptr, nullptr, nullptr, nullptr,
                             0, d_label, niter, d_memb);
            if (h_nclasses) *h_nclasses = 1;
        } else {
            const WeightInfo wi = image_extrema(*ctx, d, d_image, image_dtype);
            run_relax(*ctx, d, nadj, d_image, image_dtype, rs, d_seeds, n_seeds, d_label, niter, d_memb, h_nclasses,
                      wi);
        }
        ev_record(*ctx, 3);
        ctx->timings[2] = ev_ms(*ctx, 2, 3);
        return ROIFT_OK;
    });
}

ROIFT_API int roift_relax_binary_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                     const int32_t *d_W, const int32_t *d_Wx, const int32_t *d_Wy,
                                     const int32_t *d_Wz, const int32_t *d_seeds, int n_seeds, int32_t *d_label,
                                     int niter, float *d_memb)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = d_Wx ? 7 : nadj_of(pb);
        ROIFT_REQUIRE(image_dtype == ROIFT_I16 || image_dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16/I32");
        ev_record(*ctx, 2);
        run_relax_binary(*ctx, d, nadj, d_image, image_dtype, pb->per, d_W, d_Wx, d_Wy, d_Wz, d_seeds, n_seeds,
                         d_label, niter, d_memb);
        ev_record(*ctx, 3);
        ctx->timings[2] = ev_ms(*ctx, 2, 3);
        return ROIFT_OK;
    });
}

ROIFT_API int roift_oift_gradweight_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                        const int32_t *d_grad, float alpha, const int32_t *d_seeds, int n_seeds,
                                        int32_t *d_label, int32_t *d_cost)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        ROIFT_REQUIRE(!pb->use_per_class, ROIFT_EINVAL, "OIFT_Multi_GradWeight takes a single polarity");
        ROIFT_REQUIRE(d_grad != nullptr, ROIFT_EINVAL, "null gradient image");
        const Polarity po = polarity_of(pb);
        GradModel gm;
        gm.G = d_grad;
        gm.alpha = alpha;
        ev_record(*ctx, 0);
        const Weights w = weights_into_ctx(*ctx, d, nadj, d_image, image_dtype, po.P, &gm);
        ev_record(*ctx, 1);
        const int lv = run_oift(*ctx, d, nadj, w.wa, w.wb, w.dtype, po.lc, d_seeds, n_seeds, d_label, d_cost);
        ev_record(*ctx, 2);
        ctx->timings[0] = ev_ms(*ctx, 0, 1);
        ctx->timings[1] = ev_ms(*ctx, 1, 2);
        return lv >= 0 ? ROIFT_OK : lv;
    });
}

ROIFT_API int roift_relax_gradweight_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image,
                                         int image_dtype, const int32_t *d_grad, float alpha, const int32_t *d_seeds,
                                         int n_seeds, int32_t *d_label, int niter, float *d_memb)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        ROIFT_REQUIRE(!pb->use_per_class, ROIFT_EINVAL, "ORelax_1_Multi_GradWeight takes a single polarity");
        ROIFT_REQUIRE(d_grad != nullptr, ROIFT_EINVAL, "null gradient image");
        if (n_seeds == 0) return ROIFT_OK;   // exp_gradient_weight.cpp:296
        const RelaxSpec rs = relax_spec_of(pb);
        RelaxModel rm;
        rm.model = RELAX_GRAD;
        rm.G = d_grad;
        rm.alpha = alpha;
        ev_record(*ctx, 2);
        const WeightInfo wi = image_extrema(*ctx, d, d_image, image_dtype);
        run_relax(*ctx, d, nadj, d_image, image_dtype, rs, d_seeds, n_seeds, d_label, niter, d_memb, nullptr, wi, rm);
        ev_record(*ctx, 3);
        ctx->timings[2] = ev_ms(*ctx, 2, 3);
        return ROIFT_OK;
    });
}

ROIFT_API int roift_relax_rbf_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                  float sigma, const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter,
                                  float *d_memb, float *h_sigma)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        ROIFT_REQUIRE(!pb->use_per_class, ROIFT_EINVAL, "the RBF relaxation has no per-class form");
        RelaxModel rm;
        rm.model = RELAX_RBF;
        rm.sigma = sigma > 0.0f ? sigma : image_sigma(*ctx, d, d_image, image_dtype);
        if (h_sigma) *h_sigma = rm.sigma;
        if (n_seeds == 0) return ROIFT_OK;   // exp_gaussian_rbf_relax.cpp:235
        const RelaxSpec rs = relax_spec_of(pb);
        ev_record(*ctx, 2);
        const WeightInfo wi = image_extrema(*ctx, d, d_image, image_dtype);
        run_relax(*ctx, d, nadj, d_image, image_dtype, rs, d_seeds, n_seeds, d_label, niter, d_memb, nullptr, wi, rm);
        ev_record(*ctx, 3);
        ctx->timings[2] = ev_ms(*ctx, 2, 3);
        return ROIFT_OK;
    });
}

ROIFT_API int roift_batch_pipeline_dev(roift_ctx *ctx, int nvol, const int32_t *h_dims, float adj_radius, float per,
                                       const void *d_image, int image_dtype, const int32_t *d_seeds,
                                       const int32_t *h_nseeds, int32_t *d_label, int niter, int32_t *d_cost,
                                       int use_graph)
{
    return guarded(ctx, [&]() -> int {
        return run_batch(*ctx, nvol, h_dims, adj_radius, per, d_image, image_dtype, d_seeds, h_nseeds, d_label, niter,
                         d_cost, use_graph != 0);
    });
}

ROIFT_API int roift_batch_graph_info(const roift_ctx *c, int *captures, int *replays, int *reruns, char *err,
                                     int errlen)
{
    if (!c) return ROIFT_EINVAL;
    if (captures) *captures = c->batch_captures;
    if (replays) *replays = c->batch_replays;
    if (reruns) *reruns = c->batch_reruns;
    if (err && errlen > 0) {
        const std::string &m = c->batch_graph_error;
        const size_t k = std::min(m.size(), (size_t)errlen - 1);
        std::memcpy(err, m.data(), k);
        err[k] = 0;
    }
    return ROIFT_OK;
}

ROIFT_API int roift_relax_adjoint_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                      const int32_t *d_seeds, int n_seeds, const int32_t *d_label, int lab_obj,
                                      int niter, int ckpt_every, const float *d_scale, float sigma, float alpha,
                                      const float *d_target, float eps, double *h_loss, float *d_grad_scale,
                                      double *h_grad_sigma, double *h_grad_alpha, float *d_memb)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        ROIFT_REQUIRE(image_dtype == ROIFT_I16 || image_dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16/I32");
        run_relax_diff(*ctx, d, nadj, d_image, image_dtype, d_seeds, n_seeds, d_label, lab_obj, niter, ckpt_every,
                       d_scale, sigma, alpha, d_target, eps, h_loss, d_grad_scale, h_grad_sigma, h_grad_alpha, d_memb,
                       false);
        return ROIFT_OK;
    });
}

ROIFT_API int roift_relax_adjoint64_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image,
                                        int image_dtype, const int32_t *d_seeds, int n_seeds, const int32_t *d_label,
                                        int lab_obj, int niter, int ckpt_every, const float *d_scale, double sigma,
                                        double alpha, const float *d_target, double eps, double *h_loss,
                                        double *d_grad_scale, double *h_grad_sigma, double *h_grad_alpha,
                                        double *d_memb)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        ROIFT_REQUIRE(image_dtype == ROIFT_I16 || image_dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16/I32");
        run_relax_diff(*ctx, d, nadj, d_image, image_dtype, d_seeds, n_seeds, d_label, lab_obj, niter, ckpt_every,
                       d_scale, sigma, alpha, d_target, eps, h_loss, d_grad_scale, h_grad_sigma, h_grad_alpha, d_memb,
                       true);
        return ROIFT_OK;
    });
}

ROIFT_API int roift_pipeline_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                                 const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter, int32_t *d_cost,
                                 float *d_memb)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        const int nadj = nadj_of(pb);
        const Polarity po = polarity_of(pb);
        const RelaxSpec rs = relax_spec_of(pb);
        ev_record(*ctx, 0);
        WeightInfo wi;
        run_oift_image(*ctx, d, nadj, d_image, image_dtype, po.P, po.lc, d_seeds, n_seeds, d_label, d_cost, wi);
        ev_record(*ctx, 2);
        run_relax(*ctx, d, nadj, d_image, image_dtype, rs, d_seeds, n_seeds, d_label, niter, d_memb, nullptr, wi);
        ev_record(*ctx, 3);
        ctx->timings[0] = ctx->ktimes[0];   // the fused weight / state set-up kernel
        ctx->timings[1] = ev_ms(*ctx, 0, 2) - ctx->timings[0];
        ctx->timings[2] = ev_ms(*ctx, 2, 3);
        return ROIFT_OK;
    });
}

ROIFT_API int roift_oift_fmax(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_image, const int32_t *h_seeds,
                              int n_seeds, int32_t *h_label, int32_t *h_cost)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        int32_t *img = buf<int32_t>(*ctx, "host.img", d.n);
        int32_t *lab = buf<int32_t>(*ctx, "host.lab", d.n);
        int32_t *cost = h_cost ? buf<int32_t>(*ctx, "host.cost", d.n) : nullptr;
        int32_t *S = buf<int32_t>(*ctx, "host.seeds", n_seeds > 0 ? n_seeds : 1);
        ROIFT_HIP_CHECK(hipMemcpyAsync(img, h_image, 4 * d.n, hipMemcpyHostToDevice, ctx->stream));
        ROIFT_HIP_CHECK(hipMemcpyAsync(lab, h_label, 4 * d.n, hipMemcpyHostToDevice, ctx->stream));
        if (n_seeds > 0)
            ROIFT_HIP_CHECK(hipMemcpyAsync(S, h_seeds, 4 * (size_t)n_seeds, hipMemcpyHostToDevice, ctx->stream));
        int rc = roift_oift_fmax_dev(ctx, pb, img, ROIFT_I32, S, n_seeds, lab, cost);
        if (rc != ROIFT_OK) return rc;
        ROIFT_HIP_CHECK(hipMemcpyAsync(h_label, lab, 4 * d.n, hipMemcpyDeviceToHost, ctx->stream));
        if (h_cost) ROIFT_HIP_CHECK(hipMemcpyAsync(h_cost, cost, 4 * d.n, hipMemcpyDeviceToHost, ctx->stream));
        ROIFT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return ROIFT_OK;
    });
}

ROIFT_API int roift_oift_w(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_image, const int32_t *h_W,
                           const int32_t *h_Wx, const int32_t *h_Wy, const int32_t *h_Wz, const int32_t *h_seeds,
                           int n_seeds, int32_t *h_label, int32_t *h_cost)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        auto up = [&](const int32_t *h, const char *name) -> int32_t * {
            if (!h) return nullptr;
            int32_t *dv = buf<int32_t>(*ctx, name, d.n);
            ROIFT_HIP_CHECK(hipMemcpyAsync(dv, h, 4 * d.n, hipMemcpyHostToDevice, ctx->stream));
            return dv;
        };
        int32_t *img = up(h_image, "host.img");
        int32_t *W = up(h_W, "host.W");
        int32_t *Wx = up(h_Wx, "host.Wx");
        int32_t *Wy = up(h_Wy, "host.Wy");
        int32_t *Wz = up(h_Wz, "host.Wz");
        int32_t *lab = up(h_label, "host.lab");
        int32_t *cost = h_cost ? buf<int32_t>(*ctx, "host.cost", d.n) : nullptr;
        int32_t *S = buf<int32_t>(*ctx, "host.seeds", n_seeds > 0 ? n_seeds : 1);
        if (n_seeds > 0)
            ROIFT_HIP_CHECK(hipMemcpyAsync(S, h_seeds, 4 * (size_t)n_seeds, hipMemcpyHostToDevice, ctx->stream));
        int rc = roift_oift_w_dev(ctx, pb, img, ROIFT_I32, W, Wx, Wy, Wz, S, n_seeds, lab, cost);
        if (rc != ROIFT_OK) return rc;
        ROIFT_HIP_CHECK(hipMemcpyAsync(h_label, lab, 4 * d.n, hipMemcpyDeviceToHost, ctx->stream));
        if (h_cost) ROIFT_HIP_CHECK(hipMemcpyAsync(h_cost, cost, 4 * d.n, hipMemcpyDeviceToHost, ctx->stream));
        ROIFT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return ROIFT_OK;
    });
}

ROIFT_API int roift_relax(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_image, const int32_t *h_seeds,
                          int n_seeds, int32_t *h_label, int niter, float *h_memb)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        int32_t *img = buf<int32_t>(*ctx, "host.img", d.n);
        int32_t *lab = buf<int32_t>(*ctx, "host.lab", d.n);
        int32_t *S = buf<int32_t>(*ctx, "host.seeds", n_seeds > 0 ? n_seeds : 1);
        ROIFT_HIP_CHECK(hipMemcpyAsync(img, h_image, 4 * d.n, hipMemcpyHostToDevice, ctx->stream));
        ROIFT_HIP_CHECK(hipMemcpyAsync(lab, h_label, 4 * d.n, hipMemcpyHostToDevice, ctx->stream));
        if (n_seeds > 0)
            ROIFT_HIP_CHECK(hipMemcpyAsync(S, h_seeds, 4 * (size_t)n_seeds, hipMemcpyHostToDevice, ctx->stream));
        int32_t nc = 0;
        float *memb = nullptr;
        if (h_memb) memb = buf<float>(*ctx, "host.memb", 8 * d.n);
        int rc = roift_relax_dev(ctx, pb, img, ROIFT_I32, S, n_seeds, lab, niter, memb, &nc);
        if (rc != ROIFT_OK) return rc;
        ROIFT_HIP_CHECK(hipMemcpyAsync(h_label, lab, 4 * d.n, hipMemcpyDeviceToHost, ctx->stream));
        if (h_memb)
            ROIFT_HIP_CHECK(hipMemcpyAsync(h_memb, memb, 4 * d.n * (size_t)nc, hipMemcpyDeviceToHost, ctx->stream));
        ROIFT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return ROIFT_OK;
    });
}

ROIFT_API int roift_relax_binary(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_image, const int32_t *h_W,
                                 const int32_t *h_Wx, const int32_t *h_Wy, const int32_t *h_Wz,
                                 const int32_t *h_seeds, int n_seeds, int32_t *h_label, int niter, float *h_memb)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        ROIFT_REQUIRE(n_seeds >= 0, ROIFT_EINVAL, "negative seed count");
        auto up = [&](const int32_t *h, const char *name) -> int32_t * {
            if (!h) return nullptr;
            int32_t *p = buf<int32_t>(*ctx, name, d.n);
            ROIFT_HIP_CHECK(hipMemcpyAsync(p, h, 4 * d.n, hipMemcpyHostToDevice, ctx->stream));
            return p;
        };
        int32_t *img = up(h_image, "host.img");
        int32_t *lab = up(h_label, "host.lab");
        int32_t *W = up(h_W, "host.W"), *Wx = up(h_Wx, "host.Wx"), *Wy = up(h_Wy, "host.Wy"), *Wz = up(h_Wz, "host.Wz");
        int32_t *S = buf<int32_t>(*ctx, "host.seeds", n_seeds > 0 ? n_seeds : 1);
        if (n_seeds > 0)
            ROIFT_HIP_CHECK(hipMemcpyAsync(S, h_seeds, 4 * (size_t)n_seeds, hipMemcpyHostToDevice, ctx->stream));
        float *memb = h_memb ? buf<float>(*ctx, "host.memb", d.n) : nullptr;
        int rc = roift_relax_binary_dev(ctx, pb, img, ROIFT_I32, W, Wx, Wy, Wz, S, n_seeds, lab, niter, memb);
        if (rc != ROIFT_OK) return rc;
        ROIFT_HIP_CHECK(hipMemcpyAsync(h_label, lab, 4 * d.n, hipMemcpyDeviceToHost, ctx->stream));
        if (h_memb) ROIFT_HIP_CHECK(hipMemcpyAsync(h_memb, memb, 4 * d.n, hipMemcpyDeviceToHost, ctx->stream));
        ROIFT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return ROIFT_OK;
    });
}

ROIFT_API int roift_gaussian_blur_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_in, int image_dtype,
                                      void *d_out, int passes)
{
    return guarded(ctx, [&]() -> int {
        return run_gaussian_blur(*ctx, dims_of(pb), d_in, image_dtype, d_out, passes);
    });
}

ROIFT_API int roift_gaussian_blur(roift_ctx *ctx, const roift_problem *pb, const int32_t *h_in, int32_t *h_out,
                                  int passes)
{
    return guarded(ctx, [&]() -> int {
        const Dims d = dims_of(pb);
        ROIFT_REQUIRE(h_in && h_out, ROIFT_EINVAL, "null image pointer");
        int32_t *img = buf<int32_t>(*ctx, "host.img", d.n);
        ROIFT_HIP_CHECK(hipMemcpyAsync(img, h_in, 4 * d.n, hipMemcpyHostToDevice, ctx->stream));
        const int rc = run_gaussian_blur(*ctx, d, img, ROIFT_I32, img, passes);
        if (rc != ROIFT_OK) return rc;
        ROIFT_HIP_CHECK(hipMemcpyAsync(h_out, img, 4 * d.n, hipMemcpyDeviceToHost, ctx->stream));
        ROIFT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return ROIFT_OK;
    });
}

ROIFT_API int roift_sobel_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image, int image_dtype,
                              int32_t *d_grad, int normalize)
{
    return guarded(ctx, [&]() -> int {
        return run_sobel(*ctx, dims_of(pb), d_image, image_dtype, d_grad, normalize != 0);
    });
}

ROIFT_API int roift_bfs_edt_dev(roift_ctx *ctx, const roift_problem *pb, const int32_t *d_label,
                                const int32_t *d_seeds, int n_seeds, int32_t *d_dist)
{
    return guarded(ctx, [&]() -> int {
        ev_record(*ctx, 14);
        const int rc = run_bfs_edt(*ctx, dims_of(pb), d_label, d_seeds, n_seeds, d_dist);
        ev_record(*ctx, 15);
        ctx->ktimes[5] = ev_ms(*ctx, 14, 15);
        return rc;
    });
}

ROIFT_API int roift_boundary_seeds_dev(roift_ctx *ctx, const roift_problem *pb, int32_t *d_label, int stride,
                                       int geodesic, const int32_t *d_seeds, int n_seeds, int32_t *d_new_seeds,
                                       int capacity, int *h_added)
{
    return guarded(ctx, [&]() -> int {
        return run_boundary_seeds(*ctx, dims_of(pb), d_label, stride, geodesic != 0, d_seeds, n_seeds, d_new_seeds,
                                  capacity, h_added);
    });
}

ROIFT_API int roift_dilation_conditional_dev(roift_ctx *ctx, const roift_problem *pb, const void *d_image,
                                             int image_dtype, int32_t *d_label, float percentile, int *h_set)
{
    return guarded(ctx, [&]() -> int {
        return run_dilation_conditional(*ctx, dims_of(pb), pb->adj_radius, d_image, image_dtype, d_label, percentile,
                                        h_set);
    });
}

ROIFT_API const char *roift_io_last_error(void) { return io_last_error(); }

ROIFT_API int roift_volume_info_read(const char *path, roift_volume_info *info) { return volume_info(path, info); }

ROIFT_API int roift_volume_read(const char *path, int32_t *h_data, long long n) { return volume_read(path, h_data, n); }

ROIFT_API int roift_volume_write(const char *path, const int32_t *h_data, const roift_volume_info *info,
                                 const char *like_path)
{
    return volume_write(path, h_data, info, like_path);
}

ROIFT_API int roift_seeds_read(const char *path, int xsize, int ysize, int zsize, int32_t *h_seeds,
                               int32_t *h_labels, int capacity, int *h_count, int *h_clamped)
{
    return seeds_read(path, xsize, ysize, zsize, h_seeds, h_labels, capacity, h_count, h_clamped);
}

ROIFT_API int roift_polarity_read(const char *path, float *h_per_class, int capacity, int *h_count)
{
    return polarity_read(path, h_per_class, capacity, h_count);
}

ROIFT_API int roift_volume_read_dev(roift_ctx *ctx, const char *path, void *d_dst, int dst_dtype, int flags,
                                    int *h_shift, roift_volume_info *info)
{
    if (!ctx) return ROIFT_EINVAL;
    return guarded(ctx, [&]() -> int { return volume_read_dev(*ctx, path, d_dst, dst_dtype, flags, h_shift, info); });
}

}  // extern "C"

namespace {

// The input contract the reference's two GPU entry points share (gpu/oift_fmax_gpu.cu:171-185,
// gpu/oift_gpu.cu:192-211), over the exact OIFT_Multi_PerClass of this library:
//  - every voxel that is not a listed seed starts unlabelled, whatever h_labels holds;
//  - a seed's label is h_labels[seed], negative -> 0, and with clamp_high also > max_label -> 0;
//  - seeds outside [0, n) are skipped, and a repeated seed counts once (at its first position,
//    which is its place in the FIFO order);
//  - on return every voxel holds its label, 0 where no seed reached it (the reference unpacks
//    label 0 from the untouched (COST_INF, 0) words).
// With no valid seed, oift_gpu_run leaves h_labels as it is and returns 0 (oift_gpu.cu:207-210);
// oift_fmax_gpu_run writes label 0 everywhere and returns 0, as its untouched words unpack.
// Returns the number of level steps (the reference returns its wave count), or -1.
int legacy_run(bool gpu_run_rules, const int *h_image, int *h_labels, int n, int xsize, int ysize, int zsize,
               const float *h_per_class, int max_label, const int *h_seeds, int n_seeds)
{
    if (!h_image || !h_labels || !h_per_class || max_label < 0 || xsize <= 0 || ysize <= 0 || zsize <= 0 ||
        (long long)n != (long long)xsize * ysize * zsize || n_seeds < 0 || (n_seeds > 0 && !h_seeds))
        return -1;
    std::vector<int32_t> lab((size_t)n, -1), seeds;
    seeds.reserve((size_t)n_seeds);
    for (int i = 0; i < n_seeds; i++) {
        const int p = h_seeds[i];
        if (p < 0 || p >= n || lab[p] != -1) continue;
        int lb = h_labels[p];
        if (lb < 0) lb = 0;
        if (gpu_run_rules && lb > max_label) lb = 0;
        lab[p] = lb;
        seeds.push_back(p);
    }
    if (seeds.empty()) {
        if (!gpu_run_rules) std::fill(h_labels, h_labels + n, 0);
        return 0;
    }
    roift_ctx *ctx = nullptr;
    if (roift_create(&ctx, 0) != ROIFT_OK) return -1;
    roift_problem pb{xsize, ysize, zsize, 1.0f, 1, 0.0f, h_per_class, max_label};
    const int rc = roift_oift_fmax(ctx, &pb, h_image, seeds.data(), (int)seeds.size(), lab.data(), nullptr);
    const int steps = (int)ctx->timings[4];
    if (rc != ROIFT_OK) std::fprintf(stderr, "roift: %s\n", roift_last_error(ctx));
    roift_destroy(ctx);
    if (rc != ROIFT_OK) return -1;
    for (int p = 0; p < n; p++) h_labels[p] = lab[p] < 0 ? 0 : lab[p];
    return steps;
}

}  // namespace

extern "C" {

ROIFT_API int oift_fmax_gpu_run(const int *h_image, int *h_labels, int n, int xsize, int ysize, int zsize,
                                const float *h_per_class, int max_label, const int *h_seeds, int n_seeds)
{
    return legacy_run(false, h_image, h_labels, n, xsize, ysize, zsize, h_per_class, max_label, h_seeds, n_seeds);
}

ROIFT_API int oift_gpu_run(int cost_mode, const int *h_image, int *h_labels, int n, int xsize, int ysize, int zsize,
                           const float *h_per_class, int max_label, const int *h_seeds, int n_seeds)
{
    if (cost_mode != 0) {   // 1 = additive path cost: not part of the f_max path this library implements
        std::fprintf(stderr, "roift: oift_gpu_run cost_mode %d is not supported (f_max, 0, only)\n", cost_mode);
        return -1;
    }
    return legacy_run(true, h_image, h_labels, n, xsize, ysize, zsize, h_per_class, max_label, h_seeds, n_seeds);
}

}  // extern "C"
