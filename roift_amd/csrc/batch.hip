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

}  // namespace

int run_batch(roift_ctx &c, int nvol, const int32_t *h_dims, float adj_r, float per, const void *d_image,
              int image_dtype, const int32_t *d_seeds, const int32_t *h_nseeds, int32_t *d_label, int niter,
              int32_t *d_cost, bool use_graph)
{
    ROIFT_REQUIRE(nvol > 0 && h_dims && h_nseeds, ROIFT_EINVAL, "empty batch");
    ROIFT_REQUIRE(adj_r == 1.0f, ROIFT_EUNSUPPORTED, "the batched pipeline runs 6-adjacency (Spheric(1.0))");
    ROIFT_REQUIRE(image_dtype == ROIFT_I16 || image_dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16/I32");
    ROIFT_REQUIRE(std::fabs(per) <= 100.0f, ROIFT_EINVAL, "|per| > 100");
    hipStream_t st = c.stream;
    std::vector<PadVol> pv(nvol);
    std::vector<int> so(nvol + 1, 0);
    int PX = 0, PY = 0;
    long long off = 0, Z = 0;
    for (int v = 0; v < nvol; v++) {
        const int xs = h_dims[3 * v], ys = h_dims[3 * v + 1], zs = h_dims[3 * v + 2];
        ROIFT_REQUIRE(xs > 0 && ys > 0 && zs > 0 && h_nseeds[v] >= 0, ROIFT_EINVAL, "bad volume size or seed count");
        pv[v] = PadVol{off, xs, ys, zs, (int)Z};
        off += (long long)xs * ys * zs;
        Z += zs + 1;
        PX = std::max(PX, xs + 1);
        PY = std::max(PY, ys + 1);
        so[v + 1] = so[v] + h_nseeds[v];
    }
    const long long nflat = off, nbig = (long long)PX * PY * Z;
    ROIFT_REQUIRE(nbig < (1ll << 31), ROIFT_EUNSUPPORTED, "padded batch larger than 2^31-1 voxels");
    ROIFT_REQUIRE(Z <= 65535ll * 16, ROIFT_EUNSUPPORTED, "batch too deep");
    const int ns = so[nvol];
    PadVol *dpv = buf<PadVol>(c, "batch.pv", nvol);
    int *dso = buf<int>(c, "batch.so", nvol + 1);
    ROIFT_HIP_CHECK(hipMemcpyAsync(dpv, pv.data(), sizeof(PadVol) * nvol, hipMemcpyHostToDevice, st));
    ROIFT_HIP_CHECK(hipMemcpyAsync(dso, so.data(), sizeof(int) * (nvol + 1), hipMemcpyHostToDevice, st));
    const size_t isz = image_dtype == ROIFT_I16 ? 2 : 4;
    void *bimg = buf<uint8_t>(c, "batch.img", (size_t)nbig * isz);
    int32_t *blab = buf<int32_t>(c, "batch.lab", nbig);
    int32_t *bcost = buf<int32_t>(c, "batch.cost", nbig);
    int32_t *Sb = buf<int32_t>(c, "batch.Sb", ns > 0 ? ns : 1);
    int32_t *Sf = buf<int32_t>(c, "batch.Sf", ns > 0 ? ns : 1);
    int *info = buf<int>(c, "batch.info", 4 + 2 * nvol);
    std::vector<int> info0(4 + 2 * nvol);
    info0[0] = 0;
    info0[1] = INT32_MAX;
    info0[2] = 0;
    info0[3] = 0;
    for (int v = 0; v < nvol; v++) {
        info0[4 + 2 * v] = INT32_MIN;
        info0[5 + 2 * v] = INT32_MAX;
    }
    ROIFT_HIP_CHECK(hipMemcpyAsync(info, info0.data(), sizeof(int) * info0.size(), hipMemcpyHostToDevice, st));
    // padding: image 0, label 0 (frozen: labelled and never a seed)
    if (isz == 2) fill<int16_t>(c, (int16_t *)bimg, nbig, 0);
    else fill<int32_t>(c, (int32_t *)bimg, nbig, 0);
    fill<int32_t>(c, blab, nbig, 0);
    const dim3 g2(64, nvol);
    if (isz == 2) {
        hipLaunchKernelGGL(k_scatter<int16_t>, g2, dim3(256), 0, st, (const int16_t *)d_image, d_label, dpv, PX, PY,
                           (int16_t *)bimg, blab);
        hipLaunchKernelGGL(k_vol_extrema<int16_t>, g2, dim3(256), 0, st, (const int16_t *)d_image, dpv, info + 4);
    } else {
        hipLaunchKernelGGL(k_scatter<int32_t>, g2, dim3(256), 0, st, (const int32_t *)d_image, d_label, dpv, PX, PY,
                           (int32_t *)bimg, blab);
        hipLaunchKernelGGL(k_vol_extrema<int32_t>, g2, dim3(256), 0, st, (const int32_t *)d_image, dpv, info + 4);
    }
    if (ns > 0) {
        hipLaunchKernelGGL(k_map_seeds, dim3(8, nvol), dim3(256), 0, st, d_seeds, dso, dpv, PX, PY, Sb, Sf, info + 3);
        hipLaunchKernelGGL(k_seed_label_set, dim3(grid_for(ns)), dim3(256), 0, st, Sf, ns, d_label, info);
    }
    ROIFT_HIP_CHECK(hipGetLastError());
    std::vector<int> hinfo(info0.size());
    ROIFT_HIP_CHECK(hipMemcpyAsync(hinfo.data(), info, sizeof(int) * hinfo.size(), hipMemcpyDeviceToHost, st));
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    ROIFT_REQUIRE(hinfo[3] == 0, ROIFT_EINVAL, "seed index out of its volume");
    ROIFT_REQUIRE(hinfo[2] == 0, ROIFT_EINVAL, "seed with NIL label");
    for (int v = 0; v < nvol; v++)
        ROIFT_REQUIRE(hinfo[5 + 2 * v] >= 0, ROIFT_EINVAL, "image has negative values (shift it to >= 0 first)");

    // ---- one exact OIFT over the padded union
    const Dims bd{PX, PY, (int)Z, nbig};
    const long long slots = 3 * nbig;
    void *wa = buf<uint16_t>(c, "pipe.wa", slots), *wb = buf<uint16_t>(c, "pipe.wb", slots);
    int wdt = ROIFT_U16;
    ev_record(c, 0);
    try {
        build_weights(c, bd, 7, bimg, image_dtype, per, wa, wb, ROIFT_U16);
    } catch (const Error &e) {
        if (e.code != ROIFT_EUNSUPPORTED) throw;
        wdt = ROIFT_U32;
        wa = buf<uint32_t>(c, "pipe.wa32", slots);
        wb = buf<uint32_t>(c, "pipe.wb32", slots);
        build_weights(c, bd, 7, bimg, image_dtype, per, wa, wb, ROIFT_U32);
    }
    ev_record(c, 1);
    LabelClasses lc;
    lc.cls = {0};   // label 0 -> -per (gft_ift.cpp:968)
    lc.default_cls = 1;
    run_oift(c, bd, 7, wa, wb, wdt, lc, Sb, ns, blab, bcost);
    ev_record(c, 2);
    hipLaunchKernelGGL(k_gather, g2, dim3(256), 0, st, blab, bcost, dpv, PX, PY, d_label, d_cost);
    ROIFT_HIP_CHECK(hipGetLastError());

    // ---- relaxation over the flat volumes
    const int lab_obj = hinfo[0];
    const bool two_classes = lab_obj > 0 && hinfo[1] == lab_obj;
    if (niter > 0 && ns > 0) {
        std::vector<BatchVolume> bv(nvol);
        bool small_range = true;
        for (int v = 0; v < nvol; v++) {
            bv[v] = BatchVolume{pv[v].off, pv[v].xs, pv[v].ys, pv[v].zs, hinfo[4 + 2 * v], hinfo[5 + 2 * v]};
            small_range = small_range && hinfo[4 + 2 * v] < 0x7FFF;
        }
        if (two_classes && small_range) {
            run_relax_batch(c, bv, d_image, image_dtype, per, lab_obj, Sf, ns, d_label, niter, use_graph);
        } else {
            // general case: one relaxation per volume (each with its own class list)
            RelaxSpec rs;
            rs.per = per;
            for (int v = 0; v < nvol; v++) {
                if (h_nseeds[v] == 0) continue;
                const Dims dv{pv[v].xs, pv[v].ys, pv[v].zs, (long long)pv[v].xs * pv[v].ys * pv[v].zs};
                const void *img_v = (const uint8_t *)d_image + isz * pv[v].off;
                WeightInfo wi{hinfo[4 + 2 * v], hinfo[5 + 2 * v]};
                // the volume's seeds as volume-local indices
                run_relax(c, dv, 7, img_v, image_dtype, rs, d_seeds + so[v], h_nseeds[v], d_label + pv[v].off, niter,
                          nullptr, nullptr, wi);
            }
        }
    }
    ev_record(c, 3);
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    c.timings[0] = ev_ms(c, 0, 1);
    c.timings[1] = ev_ms(c, 1, 2);
    c.timings[2] = ev_ms(c, 2, 3);
    c.timings[3] = (double)nflat;
    return 0;
}

}  // namespace roift
