// weights.hip -- fused oriented arc-weight builder for gfx950.
//
// Replaces the weight that OIFT_Multi recomputes inside its pop loop
// (gft_ift.cpp:965-973): w = |I_p - I_q| scaled by (1 +- per/100) in double,
// truncated toward zero.  One pass reads the image once (int16 or int32) and
// writes, for every forward neighbour slot, the +per ("object") weight of both
// orientations of the edge.  The -per ("background") weight of an arc is the
// object weight of the reverse arc, so two values per undirected edge carry
// both polarities.  The same pass reduces the image extrema.
//
// Bound: HBM.  Algorithmic bytes per voxel = sizeof(image) + 2*E*sizeof(weight)
// (E = 3 forward edges for 6-adjacency, 13 for 26).
#include <map>
#include <mutex>

#include "ctx.hpp"

namespace roift {

// Block shape of the weight builder: 64 x 4 (x, y) voxels, marching WZ planes in z.
constexpr int WX = 64, WY = 4, WZ = 16;

// Base of the arc weight, to which the polarity factor (1 +- per/100) applies by the order
// of I_p and I_q (unchanged when they are equal):
// WB_IMAGE: OIFT_Multi / OIFT, |I_p-I_q|, factor in double, truncated (gft_ift.cpp:965-973).
// WB_GRAD:  OIFT_Multi_GradWeight (exp_gradient_weight.cpp:244-254), all in float:
//           wf = alpha|I_p-I_q| + (1-alpha)(G_p+G_q)*0.5, (int)(wf (1 +- per/100)), (int)wf
//           when I_p == I_q, clamped at 0.
// WB_W:     OIFT(W, A, scn, ...), the int W_p + W_q, factor in double (gft_ift.cpp:1142-1152).
// WB_WXYZ:  OIFT(Wx, Wy, Wz, scn, ...), W_t[p] + W_t[q] with t the axis of the arc
//           (gft_ift.cpp:1244-1254); Spheric(1.0) only.
// Every base is symmetric in p and q, so the two-values-per-edge layout holds for all of
// them.  This file is compiled without FMA contraction.
enum : int { WB_IMAGE = 0, WB_GRAD = 1, WB_W = 2, WB_WXYZ = 3 };
struct GradArgs {
    const int32_t *G;
    float alpha, fpf, fmf;
    const int32_t *W[3];   // WB_W: W[0]; WB_WXYZ: Wx, Wy, Wz
};

template <int NADJ, typename IT, typename WT, int BM>
__global__ __launch_bounds__(WX *WY) void k_weights(const IT *__restrict__ img, Dims d, double fp, double fm,
                                                   GradArgs ga, WT *__restrict__ wa, WT *__restrict__ wb,
                                                   int *__restrict__ ext, uint32_t *__restrict__ overflow)
{
    using A = Adjacency<NADJ>;
    constexpr uint32_t WMAXV = sizeof(WT) == 2 ? 0xFFFFu : 0xFFFFFFFFu;
    const long long xy = (long long)d.xs * d.ys;
    const int x = blockIdx.x * WX + threadIdx.x, y = blockIdx.y * WY + threadIdx.y;
    const int z0 = blockIdx.z * WZ, z1 = min(d.zs, z0 + WZ);
    int lmax = INT32_MIN, lmin = INT32_MAX;
    uint32_t ovf = 0;
    if (x < d.xs && y < d.ys) {
        for (int z = z0; z < z1; z++) {
            const long long p = x + (long long)y * d.xs + (long long)z * xy;
            const int Ip = (int)img[p];
            lmax = max(lmax, Ip);
            lmin = min(lmin, Ip);
#pragma unroll
            for (int s = 0; s < A::half; s++) {
                const int i = A::half + 1 + s;
                const int qx = x + A::dx(i), qy = y + A::dy(i), qz = z + A::dz(i);
                uint32_t a = WMAXV, b = WMAXV;
                if (in_volume(d, qx, qy, qz)) {
                    const long long q = p + A::dx(i) + (long long)A::dy(i) * d.xs + (long long)A::dz(i) * xy;
                    const int Iq = (int)img[q];
                    const int dd = abs(Ip - Iq);
                    int up, dn, eq;
                    if constexpr (BM == WB_GRAD) {
                        const float gav = (float)(ga.G[p] + ga.G[q]) * 0.5f;
                        const float wf = ga.alpha * (float)dd + (1.0f - ga.alpha) * gav;
                        eq = max(0, (int)wf);
                        up = max(0, (int)(wf * ga.fpf));
                        dn = max(0, (int)(wf * ga.fmf));
                    } else {
                        int base = dd;
                        if constexpr (BM == WB_W) base = ga.W[0][p] + ga.W[0][q];
                        if constexpr (BM == WB_WXYZ) {
                            const int t = A::dz(i) != 0 ? 2 : (A::dy(i) != 0 ? 1 : 0);
                            base = ga.W[t][p] + ga.W[t][q];
                        }
                        eq = base;
                        up = (int)((double)base * fp);
                        dn = (int)((double)base * fm);
                    }
                    // w_obj(p->q): per_pq=+per; w_obj(q->p): roles of p and q swapped
                    a = (uint32_t)(Ip > Iq ? up : (Ip < Iq ? dn : eq));
                    b = (uint32_t)(Iq > Ip ? up : (Iq < Ip ? dn : eq));
                    if (a >= WMAXV || b >= WMAXV) ovf = 1;
                }
                wa[s * d.n + p] = (WT)a;
                wb[s * d.n + p] = (WT)b;
            }
        }
    }
    // block reduction of the extrema, then one atomic per block
    __shared__ int smax[WX * WY / 64], smin[WX * WY / 64];
    for (int o = 32; o > 0; o >>= 1) {
        lmax = max(lmax, __shfl_xor(lmax, o));
        lmin = min(lmin, __shfl_xor(lmin, o));
    }
    const int tid = threadIdx.x + threadIdx.y * WX, wid = tid / 64;
    if (__lane_id() == 0) { smax[wid] = lmax; smin[wid] = lmin; }
    __syncthreads();
    if (tid == 0) {   // this workgroup's own slots: no contended atomics (k_fold_maxmin folds them)
        for (int w = 1; w < WX * WY / 64; w++) { lmax = max(lmax, smax[w]); lmin = min(lmin, smin[w]); }
        const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        ext[2 * b] = lmax;
        ext[2 * b + 1] = lmin;
    }
    if (ovf) atomicOr(overflow, 1u);
}

// ext[0] = max of part[2b], ext[1] = min of part[2b + 1] over nb workgroup partials (one
// workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void k_fold_maxmin(const int *__restrict__ part, int nb, int *__restrict__ ext)
{
    int lmax = INT32_MIN, lmin = INT32_MAX;
    for (int i = threadIdx.x; i < nb; i += 1024) {
        lmax = max(lmax, part[2 * i]);
        lmin = min(lmin, part[2 * i + 1]);
    }
    __shared__ int smax[16], smin[16];
    for (int o = 32; o > 0; o >>= 1) {
        lmax = max(lmax, __shfl_xor(lmax, o));
        lmin = min(lmin, __shfl_xor(lmin, o));
    }
    if (__lane_id() == 0) { smax[threadIdx.x / 64] = lmax; smin[threadIdx.x / 64] = lmin; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; w++) { lmax = max(lmax, smax[w]); lmin = min(lmin, smin[w]); }
        ext[0] = lmax;
        ext[1] = lmin;
    }
}

// per-workgroup extrema into ext[2b], ext[2b + 1] (folded by k_fold_maxmin)
template <typename IT>
__global__ __launch_bounds__(256) void k_extrema(const IT *__restrict__ img, long long n, int *__restrict__ ext)
{
    int lmax = INT32_MIN, lmin = INT32_MAX;
    for (long long p = gtid(); p < n; p += gstride()) {
        const int v = (int)img[p];
        lmax = max(lmax, v);
        lmin = min(lmin, v);
    }
    __shared__ int smax[8], smin[8];
    for (int o = 32; o > 0; o >>= 1) {
        lmax = max(lmax, __shfl_xor(lmax, o));
        lmin = min(lmin, __shfl_xor(lmin, o));
    }
    const int wid = threadIdx.x / 64;
    if (__lane_id() == 0) { smax[wid] = lmax; smin[wid] = lmin; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / 64); w++) { lmax = max(lmax, smax[w]); lmin = min(lmin, smin[w]); }
        ext[2 * blockIdx.x] = lmax;
        ext[2 * blockIdx.x + 1] = lmin;
    }
}

// Integer moments for estimate_sigma (exp_gaussian_rbf_relax.cpp:194-208): exact 64-bit sums.
template <typename IT>
__global__ __launch_bounds__(256) void k_moments(const IT *__restrict__ img, long long n,
                                                 unsigned long long *__restrict__ out)
{
    long long s1 = 0, s2 = 0;
    for (long long p = gtid(); p < n; p += gstride()) {
        const long long v = (long long)img[p];
        s1 += v;
        s2 += v * v;
    }
    s1 = block_reduce(s1, OpAdd());
    s2 = block_reduce(s2, OpAdd());
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], (unsigned long long)s1);
        atomicAdd(&out[1], (unsigned long long)s2);
    }
}

namespace {

dim3 weights_grid(const Dims &d)
{
    return dim3((d.xs + WX - 1) / WX, (d.ys + WY - 1) / WY, (d.zs + WZ - 1) / WZ);
}

template <int NADJ, typename IT, typename WT>
void launch_weights(roift_ctx &c, const Dims &d, const void *img, double fp, double fm, const GradArgs &ga, void *wa,
                    void *wb, int *ext, uint32_t *ovf)
{
    ROIFT_REQUIRE((d.ys + WY - 1) / WY <= 65535 && (d.zs + WZ - 1) / WZ <= 65535, ROIFT_EUNSUPPORTED,
                  "volume too large for the weight builder's grid");
    const dim3 grid = weights_grid(d);
    if (ga.G)
        hipLaunchKernelGGL((k_weights<NADJ, IT, WT, WB_GRAD>), grid, dim3(WX, WY), 0, c.stream, (const IT *)img, d,
                           fp, fm, ga, (WT *)wa, (WT *)wb, ext, ovf);
    else if (ga.W[1] && NADJ == 7)
        hipLaunchKernelGGL((k_weights<NADJ, IT, WT, WB_WXYZ>), grid, dim3(WX, WY), 0, c.stream, (const IT *)img, d,
                           fp, fm, ga, (WT *)wa, (WT *)wb, ext, ovf);
    else if (ga.W[0])
        hipLaunchKernelGGL((k_weights<NADJ, IT, WT, WB_W>), grid, dim3(WX, WY), 0, c.stream, (const IT *)img, d,
                           fp, fm, ga, (WT *)wa, (WT *)wb, ext, ovf);
    else
        hipLaunchKernelGGL((k_weights<NADJ, IT, WT, WB_IMAGE>), grid, dim3(WX, WY), 0, c.stream, (const IT *)img, d,
                           fp, fm, ga, (WT *)wa, (WT *)wb, ext, ovf);
    ROIFT_HIP_CHECK(hipGetLastError());
}

}  // namespace

// Scratch of the allocation-free builder (roift_weights_noalloc_dev): [0..1] the folded image
// extrema, [2] the overflow flag, [3] unused, then two ints (max, min) per workgroup.  Every
// call on a device shares that device's copy, so calls that overlap on one device race.
constexpr int NOALLOC_HEAD = 64;   // keeps the per-workgroup slots 256-byte aligned
static __device__ __align__(256) int scratch[NOALLOC_HEAD + 2 * NOALLOC_MAX_BLOCKS];

namespace {
std::mutex g_scratch_mu;
std::map<int, int *> g_scratch_ptr;   // device -> address of that device's copy of `scratch`
}  // namespace

int *noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) return it->second;
    // the host-side name `scratch` is not a device address: resolve this device's copy once
    // (the caller has made `device` current)
    int *p = nullptr;
    ROIFT_HIP_CHECK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    ROIFT_REQUIRE(p != nullptr, ROIFT_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = p;
    return p;
}

long long weights_blocks(const Dims &d)
{
    const dim3 wg = weights_grid(d);
    return (long long)wg.x * wg.y * wg.z;
}

void enqueue_weights(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, float per,
                     void *d_wa, void *d_wb, int weight_dtype, const GradModel *gm, int *ext, int *part, bool timed)
{
    ROIFT_REQUIRE(nadj == 7 || nadj == 27, ROIFT_EUNSUPPORTED, "adjacency must be Spheric(1.0) or Spheric(1.75)");
    ROIFT_REQUIRE(image_dtype == ROIFT_I16 || image_dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16 or I32");
    ROIFT_REQUIRE(weight_dtype == ROIFT_U16 || weight_dtype == ROIFT_U32, ROIFT_EINVAL, "weight dtype must be U16 or U32");
    // ext[0..1] are written by k_fold_maxmin; the overflow flag ext[2] starts at 0 (a fill
    // kernel, so that a captured graph holds no memset or host-copy node)
    fill<int>(c, ext, 4, 0);
    uint32_t *ovf = reinterpret_cast<uint32_t *>(ext + 2);
    const int nblk = (int)weights_blocks(d);
    // per_pq / 100.0 in double, as the reference's "w *= (1.0 +- per_pq/100.0)"
    const double fp = 1.0 + (double)per / 100.0, fm = 1.0 - (double)per / 100.0;
    GradArgs ga{nullptr, 0.0f, 0.0f, 0.0f, {nullptr, nullptr, nullptr}};
    if (gm && gm->G) {   // (1.0f +- per_pq/100.0f) in float, as the GradWeight experiment
        ga.G = gm->G;
        ga.alpha = gm->alpha;
        ga.fpf = 1.0f + per / 100.0f;
        ga.fmf = 1.0f - per / 100.0f;
    } else if (gm) {
        ROIFT_REQUIRE(!gm->Wx || nadj == 7, ROIFT_EUNSUPPORTED, "per-axis weights need Spheric(1.0)");
        ga.W[0] = gm->Wx ? gm->Wx : gm->W;
        ga.W[1] = gm->Wx ? gm->Wy : nullptr;
        ga.W[2] = gm->Wx ? gm->Wz : nullptr;
    }
    const bool i16 = image_dtype == ROIFT_I16, w16 = weight_dtype == ROIFT_U16;
    if (timed) ev_record(c, 8);
    if (nadj == 7) {
        if (i16 && w16) launch_weights<7, int16_t, uint16_t>(c, d, d_image, fp, fm, ga, d_wa, d_wb, part, ovf);
        else if (i16) launch_weights<7, int16_t, uint32_t>(c, d, d_image, fp, fm, ga, d_wa, d_wb, part, ovf);
        else if (w16) launch_weights<7, int32_t, uint16_t>(c, d, d_image, fp, fm, ga, d_wa, d_wb, part, ovf);
        else launch_weights<7, int32_t, uint32_t>(c, d, d_image, fp, fm, ga, d_wa, d_wb, part, ovf);
    } else {
        if (i16 && w16) launch_weights<27, int16_t, uint16_t>(c, d, d_image, fp, fm, ga, d_wa, d_wb, part, ovf);
        else if (i16) launch_weights<27, int16_t, uint32_t>(c, d, d_image, fp, fm, ga, d_wa, d_wb, part, ovf);
        else if (w16) launch_weights<27, int32_t, uint16_t>(c, d, d_image, fp, fm, ga, d_wa, d_wb, part, ovf);
        else launch_weights<27, int32_t, uint32_t>(c, d, d_image, fp, fm, ga, d_wa, d_wb, part, ovf);
    }
    hipLaunchKernelGGL(k_fold_maxmin, dim3(1), dim3(1024), 0, c.stream, (const int *)part, nblk, ext);
    ROIFT_HIP_CHECK(hipGetLastError());
    if (timed) ev_record(c, 9);
}

WeightInfo build_weights(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, float per,
                         void *d_wa, void *d_wb, int weight_dtype, const GradModel *gm, int *d_scratch)
{
    const long long nblk_ll = weights_blocks(d);
    int *ext, *part;
    if (d_scratch) {
        // caller-provided scratch (the library's static array): sized for NOALLOC_MAX_BLOCKS
        // workgroups, cleared by a fill kernel (no memset / host copy nodes)
        ROIFT_REQUIRE(nblk_ll <= NOALLOC_MAX_BLOCKS, ROIFT_EUNSUPPORTED,
                      "volume needs " + std::to_string(nblk_ll) + " weight-builder workgroups; the allocation-free "
                      "entry holds scratch for at most " + std::to_string(NOALLOC_MAX_BLOCKS) +
                      " (ceil(x/64) * ceil(y/4) * ceil(z/16))");
        ext = d_scratch;
        part = d_scratch + NOALLOC_HEAD;
    } else {
        ext = buf<int>(c, "w.ext", 4);
        part = buf<int>(c, "w.part", 2 * (size_t)nblk_ll);
    }
    enqueue_weights(c, d, nadj, d_image, image_dtype, per, d_wa, d_wb, weight_dtype, gm, ext, part, true);
    int host[4];
    ROIFT_HIP_CHECK(hipMemcpyAsync(host, ext, sizeof(host), hipMemcpyDeviceToHost, c.stream));
    ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
    c.ktimes[0] = ev_ms(c, 8, 9);
    ROIFT_REQUIRE(host[2] == 0, ROIFT_EUNSUPPORTED, "arc weight exceeds the weight dtype range (use U32 weights)");
    return WeightInfo{host[0], host[1]};
}

WeightInfo image_extrema(roift_ctx &c, const Dims &d, const void *d_image, int image_dtype)
{
    int *ext = buf<int>(c, "w.ext", 4);
    const int init[4] = {INT32_MIN, INT32_MAX, 0, 0};
    ROIFT_HIP_CHECK(hipMemcpyAsync(ext, init, sizeof(init), hipMemcpyHostToDevice, c.stream));
    const unsigned g = grid_for(d.n, 256, 4096);
    int *part = buf<int>(c, "w.part", 2 * (size_t)g);
    if (image_dtype == ROIFT_I16)
        hipLaunchKernelGGL(k_extrema<int16_t>, dim3(g), dim3(256), 0, c.stream, (const int16_t *)d_image, d.n, part);
    else
        hipLaunchKernelGGL(k_extrema<int32_t>, dim3(g), dim3(256), 0, c.stream, (const int32_t *)d_image, d.n, part);
    ROIFT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_fold_maxmin, dim3(1), dim3(1024), 0, c.stream, (const int *)part, (int)g, ext);
    ROIFT_HIP_CHECK(hipGetLastError());
    int host[4];
    ROIFT_HIP_CHECK(hipMemcpyAsync(host, ext, sizeof(host), hipMemcpyDeviceToHost, c.stream));
    ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
    return WeightInfo{host[0], host[1]};
}

}  // namespace roift

namespace roift {

float image_sigma(roift_ctx &c, const Dims &d, const void *d_image, int image_dtype)
{
    unsigned long long *m = buf<unsigned long long>(c, "w.moments", 2);
    const unsigned long long zero[2] = {0, 0};
    ROIFT_HIP_CHECK(hipMemcpyAsync(m, zero, sizeof(zero), hipMemcpyHostToDevice, c.stream));
    if (image_dtype == ROIFT_I16)
        hipLaunchKernelGGL(k_moments<int16_t>, dim3(grid_for(d.n, 256, 1024)), dim3(256), 0, c.stream,
                           (const int16_t *)d_image, d.n, m);
    else
        hipLaunchKernelGGL(k_moments<int32_t>, dim3(grid_for(d.n, 256, 1024)), dim3(256), 0, c.stream,
                           (const int32_t *)d_image, d.n, m);
    ROIFT_HIP_CHECK(hipGetLastError());
    unsigned long long h[2];
    ROIFT_HIP_CHECK(hipMemcpyAsync(h, m, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
    const double mean = (double)(long long)h[0] / (double)d.n;
    double var = (double)(long long)h[1] / (double)d.n - mean * mean;
    if (var < 1.0) var = 1.0;
    return (float)std::sqrt(var);
}

}  // namespace roift
