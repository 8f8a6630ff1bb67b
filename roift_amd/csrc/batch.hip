// batch.hip -- OIFT_Multi + ORelax_1_Multi over a batch of independent volumes.
//
// The volumes share nothing, so one exact OIFT over their disjoint union gives every
// volume its own forest: levels are global, but a volume with no entry at a level simply
// takes no part in it, each volume's FIFO keys keep their relative order, and the
// competition bound r_safe is only ever more conservative.  The union is laid out as one
// box of PX x PY x Z voxels (PX = max xs + 1, PY = max ys + 1) with the volumes stacked in
// z, one separator plane after each; voxels outside a volume's own box are frozen
// (labelled, never seeds), so no arc leaves a volume.  One persistent OIFT launch then
// serves the whole batch.  The relaxation runs on the flat (unpadded) volumes with one
// launch per iteration for all of them (run_relax_batch), replayed as a HIP graph.
#include <algorithm>

#include "ctx.hpp"

namespace roift {
namespace {

struct PadVol {
    long long off;   // flat offset
    int xs, ys, zs, zb;   // size, first plane in the padded box
};

template <typename IT>
__global__ void k_scatter(const IT *__restrict__ img, const int32_t *__restrict__ lab, const PadVol *__restrict__ pv,
                          int PX, int PY, IT *__restrict__ bimg, int32_t *__restrict__ blab)
{
    const PadVol v = pv[blockIdx.y];
    const long long n = (long long)v.xs * v.ys * v.zs;
    const long long pxy = (long long)PX * PY;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % v.xs), y = (int)((i / v.xs) % v.ys), z = (int)(i / ((long long)v.xs * v.ys));
        const long long P = x + (long long)y * PX + (long long)(v.zb + z) * pxy;
        bimg[P] = img[v.off + i];
        blab[P] = lab[v.off + i];
    }
}

__global__ void k_gather(const int32_t *__restrict__ blab, const int32_t *__restrict__ bcost,
                         const PadVol *__restrict__ pv, int PX, int PY, int32_t *__restrict__ lab,
                         int32_t *__restrict__ cost)
{
    const PadVol v = pv[blockIdx.y];
    const long long n = (long long)v.xs * v.ys * v.zs;
    const long long pxy = (long long)PX * PY;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % v.xs), y = (int)((i / v.xs) % v.ys), z = (int)(i / ((long long)v.xs * v.ys));
        const long long P = x + (long long)y * PX + (long long)(v.zb + z) * pxy;
        lab[v.off + i] = blab[P];
        if (cost) cost[v.off + i] = bcost[P];
    }
}

// Volume-local seed indices -> padded and flat indices (volume v's seeds are S[so[v] .. so[v+1])).
__global__ void k_map_seeds(const int32_t *__restrict__ S, const int *__restrict__ so, const PadVol *__restrict__ pv,
                            int PX, int PY, int32_t *__restrict__ Sb, int32_t *__restrict__ Sf, int *__restrict__ bad)
{
    const int vi = blockIdx.y;
    const PadVol v = pv[vi];
    const long long n = (long long)v.xs * v.ys * v.zs;
    for (int j = so[vi] + (int)(blockIdx.x * blockDim.x + threadIdx.x); j < so[vi + 1]; j += gridDim.x * blockDim.x) {
        const long long i = S[j];
        if (i < 0 || i >= n) {
            atomicOr(bad, 1);
            Sb[j] = -1;
            Sf[j] = -1;
            continue;
        }
        const int x = (int)(i % v.xs), y = (int)((i / v.xs) % v.ys), z = (int)(i / ((long long)v.xs * v.ys));
        Sb[j] = (int32_t)(x + (long long)y * PX + (long long)(v.zb + z) * PX * PY);
        Sf[j] = (int32_t)(v.off + i);
    }
}

// Per-volume image extrema (blockIdx.y = volume).
template <typename IT>
__global__ __launch_bounds__(256) void k_vol_extrema(const IT *__restrict__ img, const PadVol *__restrict__ pv,
                                                     int *__restrict__ ext)
{
    const PadVol v = pv[blockIdx.y];
    const long long n = (long long)v.xs * v.ys * v.zs;
    int lmax = INT32_MIN, lmin = INT32_MAX;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int val = (int)img[v.off + i];
        lmax = max(lmax, val);
        lmin = min(lmin, val);
    }
    lmax = block_reduce(lmax, OpMax());
    lmin = block_reduce(lmin, OpMin());
    if (threadIdx.x == 0) {
        atomicMax(&ext[2 * blockIdx.y], lmax);
        atomicMin(&ext[2 * blockIdx.y + 1], lmin);
    }
}

// Seed labels of the whole batch, for the class check (labels must be {0, L}).
__global__ void k_seed_label_set(const int32_t *__restrict__ Sf, int ns, const int32_t *__restrict__ lab,
                                 int *__restrict__ info)
{
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int s = Sf[i];
        if (s < 0) continue;
        const int lb = lab[s];
        if (lb > 0) {
            atomicMax(&info[0], lb);
            atomicMin(&info[1], lb);
        }
        if (lb < 0) atomicOr(&info[2], 1);
    }
}

// info[0] = largest and info[1] = smallest positive seed label, info[2] = a seed with NIL label,
// info[3] = a seed outside its volume, info[4 + 2v] / info[5 + 2v] = volume v's image max / min.
__global__ void k_batch_info_init(int *__restrict__ info, int nvol)
{
    for (int i = (int)gtid(); i < 4 + 2 * nvol; i += (int)gstride())
        info[i] = i == 1 ? INT32_MAX : (i < 4 ? 0 : ((i & 1) ? INT32_MAX : INT32_MIN));
}

__global__ void k_copy_i32(const int32_t *__restrict__ src, int32_t *__restrict__ dst, long long n)
{
    for (long long i = gtid(); i < n; i += gstride()) dst[i] = src[i];
}

// The padded union's layout: shapes only (no device data).
struct Layout {
    std::vector<PadVol> pv;
    std::vector<int> so;
    std::vector<BatchVolume> bv;   // the flat volumes (extrema unused: the relaxation reads them on the device)
    int PX = 0, PY = 0;
    long long nflat = 0, nbig = 0, Z = 0;
    int ns = 0;
};

Layout layout_of(int nvol, const int32_t *h_dims, const int32_t *h_nseeds)
{
    Layout L;
    L.pv.resize(nvol);
    L.so.assign(nvol + 1, 0);
    long long off = 0;
    for (int v = 0; v < nvol; v++) {
        const int xs = h_dims[3 * v], ys = h_dims[3 * v + 1], zs = h_dims[3 * v + 2];
        ROIFT_REQUIRE(xs > 0 && ys > 0 && zs > 0 && h_nseeds[v] >= 0, ROIFT_EINVAL, "bad volume size or seed count");
        L.pv[v] = PadVol{off, xs, ys, zs, (int)L.Z};
        L.bv.push_back(BatchVolume{off, xs, ys, zs, 0, 0});
        off += (long long)xs * ys * zs;
        L.Z += zs + 1;
        L.PX = std::max(L.PX, xs + 1);
        L.PY = std::max(L.PY, ys + 1);
        L.so[v + 1] = L.so[v] + h_nseeds[v];
    }
    L.nflat = off;
    L.nbig = (long long)L.PX * L.PY * L.Z;
    ROIFT_REQUIRE(L.nbig < (1ll << 31), ROIFT_EUNSUPPORTED, "padded batch larger than 2^31-1 voxels");
    ROIFT_REQUIRE(L.Z <= 65535ll * 16, ROIFT_EUNSUPPORTED, "batch too deep");
    L.ns = L.so[nvol];
    return L;
}

struct BatchBufs {
    PadVol *dpv;
    int *dso;
    void *bimg;
    int32_t *blab, *bcost, *Sb, *Sf;
    int *info;
    void *wa, *wb;   // planar 16-bit weights
};

BatchBufs batch_bufs(roift_ctx &c, const Layout &L, int nvol, int image_dtype)
{
    BatchBufs b;
    b.dpv = buf<PadVol>(c, "batch.pv", nvol);
    b.dso = buf<int>(c, "batch.so", nvol + 1);
    const size_t isz = image_dtype == ROIFT_I16 ? 2 : 4;
    b.bimg = buf<uint8_t>(c, "batch.img", (size_t)L.nbig * isz);
    b.blab = buf<int32_t>(c, "batch.lab", L.nbig);
    b.bcost = buf<int32_t>(c, "batch.cost", L.nbig);
    b.Sb = buf<int32_t>(c, "batch.Sb", L.ns > 0 ? L.ns : 1);
    b.Sf = buf<int32_t>(c, "batch.Sf", L.ns > 0 ? L.ns : 1);
    b.info = buf<int>(c, "batch.info", 4 + 2 * nvol);
    b.wa = buf<uint16_t>(c, "pipe.wa", 3 * L.nbig);
    b.wb = buf<uint16_t>(c, "pipe.wb", 3 * L.nbig);
    return b;
}

// The union box from the flat volumes and the seed tables, with the per-volume checks'
// inputs (info); launches only.
void enqueue_front(roift_ctx &c, const Layout &L, const BatchBufs &b, int nvol, const void *d_image, int image_dtype,
                   const int32_t *d_seeds, const int32_t *d_label)
{
    hipStream_t st = c.stream;
    hipLaunchKernelGGL(k_batch_info_init, dim3(grid_for(4 + 2 * nvol)), dim3(256), 0, st, b.info, nvol);
    // padding: image 0, label 0 (frozen: labelled and never a seed)
    if (image_dtype == ROIFT_I16) fill<int16_t>(c, (int16_t *)b.bimg, L.nbig, 0);
    else fill<int32_t>(c, (int32_t *)b.bimg, L.nbig, 0);
    fill<int32_t>(c, b.blab, L.nbig, 0);
    const dim3 g2(64, nvol);
    if (image_dtype == ROIFT_I16) {
        hipLaunchKernelGGL(k_scatter<int16_t>, g2, dim3(256), 0, st, (const int16_t *)d_image, d_label, b.dpv, L.PX,
                           L.PY, (int16_t *)b.bimg, b.blab);
        hipLaunchKernelGGL(k_vol_extrema<int16_t>, g2, dim3(256), 0, st, (const int16_t *)d_image, b.dpv, b.info + 4);
    } else {
        hipLaunchKernelGGL(k_scatter<int32_t>, g2, dim3(256), 0, st, (const int32_t *)d_image, d_label, b.dpv, L.PX,
                           L.PY, (int32_t *)b.bimg, b.blab);
        hipLaunchKernelGGL(k_vol_extrema<int32_t>, g2, dim3(256), 0, st, (const int32_t *)d_image, b.dpv, b.info + 4);
    }
    if (L.ns > 0) {
        hipLaunchKernelGGL(k_map_seeds, dim3(8, nvol), dim3(256), 0, st, d_seeds, b.dso, b.dpv, L.PX, L.PY, b.Sb, b.Sf,
                           b.info + 3);
        hipLaunchKernelGGL(k_seed_label_set, dim3(grid_for(L.ns)), dim3(256), 0, st, b.Sf, L.ns, d_label, b.info);
    }
    ROIFT_HIP_CHECK(hipGetLastError());
}

// The per-volume checks, in the order the eager pipeline makes them.
void check_info(const int *hinfo, int nvol)
{
    ROIFT_REQUIRE(hinfo[3] == 0, ROIFT_EINVAL, "seed index out of its volume");
    ROIFT_REQUIRE(hinfo[2] == 0, ROIFT_EINVAL, "seed with NIL label");
    for (int v = 0; v < nvol; v++)
        ROIFT_REQUIRE(hinfo[5 + 2 * v] >= 0, ROIFT_EINVAL, "image has negative values (shift it to >= 0 first)");
}

// The batch's relaxation runs as one batched launch per iteration when its seeds hold labels
// {0, L} and every intensity is below 32767; otherwise one relaxation per volume.
bool batched_relax_applies(const int *hinfo, int nvol)
{
    const int lab_obj = hinfo[0];
    bool ok = lab_obj > 0 && hinfo[1] == lab_obj;
    for (int v = 0; v < nvol; v++) ok = ok && hinfo[4 + 2 * v] < 0x7FFF;
    return ok;
}

LabelClasses batch_classes()
{
    LabelClasses lc;
    lc.cls = {0};   // label 0 -> -per (gft_ift.cpp:968)
    lc.default_cls = 1;
    return lc;
}

int run_batch_eager(roift_ctx &c, const Layout &L, int nvol, float per, const void *d_image, int image_dtype,
                    const int32_t *d_seeds, const int32_t *h_nseeds, int32_t *d_label, int niter, int32_t *d_cost)
{
    hipStream_t st = c.stream;
    const BatchBufs b = batch_bufs(c, L, nvol, image_dtype);
    ROIFT_HIP_CHECK(hipMemcpyAsync(b.dpv, L.pv.data(), sizeof(PadVol) * nvol, hipMemcpyHostToDevice, st));
    ROIFT_HIP_CHECK(hipMemcpyAsync(b.dso, L.so.data(), sizeof(int) * (nvol + 1), hipMemcpyHostToDevice, st));
    enqueue_front(c, L, b, nvol, d_image, image_dtype, d_seeds, d_label);
    std::vector<int> hinfo(4 + 2 * nvol);
    ROIFT_HIP_CHECK(hipMemcpyAsync(hinfo.data(), b.info, sizeof(int) * hinfo.size(), hipMemcpyDeviceToHost, st));
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    check_info(hinfo.data(), nvol);

    // ---- one exact OIFT over the padded union
    const Dims bd{L.PX, L.PY, (int)L.Z, L.nbig};
    void *wa = b.wa, *wb = b.wb;
    int wdt = ROIFT_U16;
    ev_record(c, 0);
    try {
        build_weights(c, bd, 7, b.bimg, image_dtype, per, wa, wb, ROIFT_U16);
    } catch (const Error &e) {
        if (e.code != ROIFT_EUNSUPPORTED) throw;
        wdt = ROIFT_U32;
        wa = buf<uint32_t>(c, "pipe.wa32", 3 * L.nbig);
        wb = buf<uint32_t>(c, "pipe.wb32", 3 * L.nbig);
        build_weights(c, bd, 7, b.bimg, image_dtype, per, wa, wb, ROIFT_U32);
    }
    ev_record(c, 1);
    run_oift(c, bd, 7, wa, wb, wdt, batch_classes(), b.Sb, L.ns, b.blab, b.bcost);
    ev_record(c, 2);
    hipLaunchKernelGGL(k_gather, dim3(64, nvol), dim3(256), 0, st, b.blab, b.bcost, b.dpv, L.PX, L.PY, d_label, d_cost);
    ROIFT_HIP_CHECK(hipGetLastError());

    // ---- relaxation over the flat volumes
    if (niter > 0 && L.ns > 0) {
        if (batched_relax_applies(hinfo.data(), nvol)) {
            run_relax_batch(c, L.bv, d_image, image_dtype, per, b.info, b.info + 4, b.Sf, L.ns, d_label, niter,
                            RELAX_BATCH_EAGER);
        } else {
            // general case: one relaxation per volume (each with its own class list)
            RelaxSpec rs;
            rs.per = per;
            const size_t isz = image_dtype == ROIFT_I16 ? 2 : 4;
            for (int v = 0; v < nvol; v++) {
                if (h_nseeds[v] == 0) continue;
                const PadVol &pv = L.pv[v];
                const Dims dv{pv.xs, pv.ys, pv.zs, (long long)pv.xs * pv.ys * pv.zs};
                const void *img_v = (const uint8_t *)d_image + isz * pv.off;
                WeightInfo wi{hinfo[4 + 2 * v], hinfo[5 + 2 * v]};
                // the volume's seeds as volume-local indices
                run_relax(c, dv, 7, img_v, image_dtype, rs, d_seeds + L.so[v], h_nseeds[v], d_label + pv.off, niter,
                          nullptr, nullptr, wi);
            }
        }
    }
    ev_record(c, 3);
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    c.timings[0] = ev_ms(c, 0, 1);
    c.timings[1] = ev_ms(c, 1, 2);
    c.timings[2] = ev_ms(c, 2, 3);
    c.timings[3] = (double)L.nflat;
    return 0;
}

// Everything run_batch_eager does for its common case, as launches only (u16 weights, the batched
// relaxation), then the status words copied to pinned host memory: h_bstat[0 .. 4 + 2 nvol) =
// info, then the weight builder's image max, min and overflow flag; the OIFT's counters go to the
// pinned counter mirror.  The input labels are saved first (batch.lab0) for a rerun.
void enqueue_graph_body(roift_ctx &c, const Layout &L, const BatchBufs &b, int nvol, float per, const void *d_image,
                        int image_dtype, const int32_t *d_seeds, int32_t *d_label, int niter, int32_t *d_cost,
                        int32_t *lab0, int *wext, int *wpart)
{
    hipStream_t st = c.stream;
    hipLaunchKernelGGL(k_copy_i32, dim3(grid_for(L.nflat, 256, 8192)), dim3(256), 0, st, (const int32_t *)d_label, lab0,
                       L.nflat);
    enqueue_front(c, L, b, nvol, d_image, image_dtype, d_seeds, d_label);
    const Dims bd{L.PX, L.PY, (int)L.Z, L.nbig};
    enqueue_weights(c, bd, 7, b.bimg, image_dtype, per, b.wa, b.wb, ROIFT_U16, nullptr, wext, wpart, false);
    run_oift_staged(c, OIFT_ENQUEUE, bd, 7, b.wa, b.wb, ROIFT_U16, batch_classes(), b.Sb, L.ns, b.blab, b.bcost);
    hipLaunchKernelGGL(k_gather, dim3(64, nvol), dim3(256), 0, st, b.blab, b.bcost, b.dpv, L.PX, L.PY, d_label, d_cost);
    ROIFT_HIP_CHECK(hipGetLastError());
    if (niter > 0)
        run_relax_batch(c, L.bv, d_image, image_dtype, per, b.info, b.info + 4, b.Sf, L.ns, d_label, niter,
                        RELAX_BATCH_CAPTURE);
    const int ninfo = 4 + 2 * nvol;
    ROIFT_HIP_CHECK(hipMemcpyAsync(c.h_bstat, b.info, sizeof(int) * ninfo, hipMemcpyDeviceToHost, st));
    ROIFT_HIP_CHECK(hipMemcpyAsync(c.h_bstat + ninfo, wext, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
}

std::string graph_key(roift_ctx &c, int nvol, const int32_t *h_dims, const int32_t *h_nseeds, float per,
                      const void *d_image, int image_dtype, const int32_t *d_seeds, const int32_t *d_label, int niter,
                      const int32_t *d_cost)
{
    std::string k;
    char t[256];
    std::snprintf(t, sizeof(t), "%d %a %p %d %p %p %d %p | %d %d %d |", nvol, (double)per, d_image, image_dtype,
                  (const void *)d_seeds, (const void *)d_label, niter, (const void *)d_cost, c.tune_blocks,
                  c.tune_budget, c.tune_flood_mode);
    k += t;
    for (int v = 0; v < nvol; v++) {
        std::snprintf(t, sizeof(t), "%d,%d,%d,%d;", h_dims[3 * v], h_dims[3 * v + 1], h_dims[3 * v + 2], h_nseeds[v]);
        k += t;
    }
    return k;
}

}  // namespace

int run_batch(roift_ctx &c, int nvol, const int32_t *h_dims, float adj_r, float per, const void *d_image,
              int image_dtype, const int32_t *d_seeds, const int32_t *h_nseeds, int32_t *d_label, int niter,
              int32_t *d_cost, bool use_graph)
{
    ROIFT_REQUIRE(nvol > 0 && h_dims && h_nseeds, ROIFT_EINVAL, "empty batch");
    ROIFT_REQUIRE(adj_r == 1.0f, ROIFT_EUNSUPPORTED, "the batched pipeline runs 6-adjacency (Spheric(1.0))");
    ROIFT_REQUIRE(image_dtype == ROIFT_I16 || image_dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16/I32");
    ROIFT_REQUIRE(std::fabs(per) <= 100.0f, ROIFT_EINVAL, "|per| > 100");
    ROIFT_REQUIRE(niter >= 0, ROIFT_EINVAL, "negative iteration count");
    const Layout L = layout_of(nvol, h_dims, h_nseeds);
    if (!use_graph || L.ns == 0)
        return run_batch_eager(c, L, nvol, per, d_image, image_dtype, d_seeds, h_nseeds, d_label, niter, d_cost);

    hipStream_t st = c.stream;
    const std::string key =
        graph_key(c, nvol, h_dims, h_nseeds, per, d_image, image_dtype, d_seeds, d_label, niter, d_cost);
    if (key == c.batch_graph_fail_key)   // this key's capture failed before: eager
        return run_batch_eager(c, L, nvol, per, d_image, image_dtype, d_seeds, h_nseeds, d_label, niter, d_cost);
    if (!c.batch_graph || c.batch_graph_key != key || c.batch_graph_gen != c.buf_gen) {
        // A new key: the eager pipeline runs this call (and allocates every buffer the graph
        // uses), then the graph is captured for the calls that follow.
        if (c.batch_graph) {
            ROIFT_HIP_CHECK(hipGraphExecDestroy(c.batch_graph));
            c.batch_graph = nullptr;
        }
        const int r = run_batch_eager(c, L, nvol, per, d_image, image_dtype, d_seeds, h_nseeds, d_label, niter, d_cost);
        // buffers only the graph uses, the pinned status words, the OIFT's buffers and class table
        const BatchBufs b = batch_bufs(c, L, nvol, image_dtype);
        (void)buf<int32_t>(c, "batch.lab0", L.nflat);
        (void)buf<int>(c, "batch.wext", 4);
        (void)buf<int>(c, "batch.wpart", 2 * weights_blocks(Dims{L.PX, L.PY, (int)L.Z, L.nbig}));
        const Dims bd{L.PX, L.PY, (int)L.Z, L.nbig};
        run_oift_staged(c, OIFT_PREPARE, bd, 7, b.wa, b.wb, ROIFT_U16, batch_classes(), b.Sb, L.ns, b.blab, b.bcost);
        const size_t nstat = 4 + 2 * (size_t)nvol + 4;
        if (c.h_bstat_n < nstat) {
            if (c.h_bstat) ROIFT_HIP_CHECK(hipHostFree(c.h_bstat));
            c.h_bstat = nullptr;
            c.h_bstat_n = 0;
            ROIFT_HIP_CHECK(hipHostMalloc((void **)&c.h_bstat, sizeof(int) * nstat, hipHostMallocDefault));
            c.h_bstat_n = nstat;
        }
        // the shape tables (also uploaded before every replay: another batch may overwrite them)
        ROIFT_HIP_CHECK(hipMemcpyAsync(b.dpv, L.pv.data(), sizeof(PadVol) * nvol, hipMemcpyHostToDevice, st));
        ROIFT_HIP_CHECK(hipMemcpyAsync(b.dso, L.so.data(), sizeof(int) * (nvol + 1), hipMemcpyHostToDevice, st));
        if (niter > 0)   // the relaxation's tile and volume tables
            run_relax_batch(c, L.bv, d_image, image_dtype, per, b.info, b.info + 4, b.Sf, L.ns, d_label, niter,
                            RELAX_BATCH_TABLES);
        ROIFT_HIP_CHECK(hipStreamSynchronize(st));
        const unsigned long long gen = c.buf_gen;
        int32_t *lab0 = buf<int32_t>(c, "batch.lab0", L.nflat);
        int *wext = buf<int>(c, "batch.wext", 4);
        int *wpart = buf<int>(c, "batch.wpart", 2 * weights_blocks(bd));
        hipGraph_t g = nullptr;
        try {
            ROIFT_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            try {
                enqueue_graph_body(c, L, b, nvol, per, d_image, image_dtype, d_seeds, d_label, niter, d_cost, lab0,
                                   wext, wpart);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            ROIFT_HIP_CHECK(hipStreamEndCapture(st, &g));
            ROIFT_REQUIRE(c.buf_gen == gen, ROIFT_EHIP, "a buffer moved while the batch graph was captured");
            const hipError_t ie = hipGraphInstantiate(&c.batch_graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            ROIFT_HIP_CHECK(ie);
        } catch (const Error &e) {
            // the runtime could not record this pipeline: the key runs eagerly from now on
            (void)hipGetLastError();
            c.batch_graph = nullptr;
            c.batch_graph_error = e.msg;
            c.batch_graph_fail_key = key;
            return r;
        }
        c.batch_graph_key = key;
        c.batch_graph_gen = gen;
        c.batch_captures++;
        return r;
    }

    // ---- replay.  The tables the graph reads but does not write (shapes, the OIFT's label classes)
    // are uploaded again first: other calls on this context share those buffers.
    const BatchBufs b = batch_bufs(c, L, nvol, image_dtype);   // cached: the captured addresses
    ROIFT_HIP_CHECK(hipMemcpyAsync(b.dpv, L.pv.data(), sizeof(PadVol) * nvol, hipMemcpyHostToDevice, st));
    ROIFT_HIP_CHECK(hipMemcpyAsync(b.dso, L.so.data(), sizeof(int) * (nvol + 1), hipMemcpyHostToDevice, st));
    run_oift_staged(c, OIFT_PREPARE, Dims{L.PX, L.PY, (int)L.Z, L.nbig}, 7, b.wa, b.wb, ROIFT_U16, batch_classes(),
                    b.Sb, L.ns, b.blab, b.bcost);
    if (niter > 0)
        run_relax_batch(c, L.bv, d_image, image_dtype, per, b.info, b.info + 4, b.Sf, L.ns, d_label, niter,
                            RELAX_BATCH_TABLES);
    ROIFT_REQUIRE(c.batch_graph_gen == c.buf_gen, ROIFT_EHIP, "batch graph buffers moved");
    ev_record(c, 0);
    ROIFT_HIP_CHECK(hipGraphLaunch(c.batch_graph, st));
    ev_record(c, 3);
    c.batch_replays++;
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    const int *hinfo = c.h_bstat;
    const int *hw = c.h_bstat + 4 + 2 * nvol;
    int32_t *lab0 = buf<int32_t>(c, "batch.lab0", L.nflat);
    auto restore = [&]() {   // the caller's labels as they were (errors and reruns leave no partial result)
        hipLaunchKernelGGL(k_copy_i32, dim3(grid_for(L.nflat, 256, 8192)), dim3(256), 0, st, (const int32_t *)lab0,
                           d_label, L.nflat);
        ROIFT_HIP_CHECK(hipGetLastError());
        ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    };
    try {
        check_info(hinfo, nvol);
        const Dims bd{L.PX, L.PY, (int)L.Z, L.nbig};
        if (hw[2] == 0)   // the 16-bit weights held: the OIFT ran on them
            run_oift_staged(c, OIFT_FINISH, bd, 7, b.wa, b.wb, ROIFT_U16, batch_classes(), b.Sb, L.ns, b.blab, b.bcost);
    } catch (...) {
        restore();
        throw;
    }
    if (hw[2] != 0 || (niter > 0 && !batched_relax_applies(hinfo, nvol))) {
        // weights beyond 16 bits or a relaxation the batched launch does not cover: the eager
        // pipeline reruns the call from the saved labels
        restore();
        c.batch_reruns++;
        return run_batch_eager(c, L, nvol, per, d_image, image_dtype, d_seeds, h_nseeds, d_label, niter, d_cost);
    }
    c.timings[0] = 0.0;   // one graph: stage times are not split
    c.timings[1] = 0.0;
    c.timings[2] = ev_ms(c, 0, 3);
    c.timings[3] = (double)L.nflat;
    return 0;
}

}  // namespace roift
