// adjoint.hip -- differentiable two-class relaxation and its adjoint (reverse mode with
// checkpoints) for gfx950.
//
// The learnable relaxation weight has no reference counterpart: it joins the Gaussian-RBF
// weight of ORelax_1_Multi_GaussRBF (exp_gaussian_rbf_relax.cpp:343-345) with the soft
// polarity of ORelax_1_Multi (gft_ift.cpp:8800-8812) and a per-voxel scale:
//   w_pq = scale_q * exp(-(I_p - I_q)^2 / sigma^2) * (1 + alpha sgn(I_p - I_q)(2 f_q - 1)) / Dpq
// f = background membership; f'_p = sum w f_q / sum w for non-seed p (f_p when the sum is
// 0); seeds keep their value.  Loss: BCE of o = 1 - f against a target (o clamped to
// [eps, 1 - eps]).  The backward pass stores one state every `ckpt` iterations, recomputes
// each segment forward, and runs two passes per iteration:
//   pass 1 (per p):  B_p = sum_q w_pq,  c_p = lambda'_p / B_p  (the self term when B_p = 0)
//   pass 2 (per q):  lambda_q = sum_{p ~ q} c_p (w_pq + (f_q - f'_p) dw_pq/df_q)
//                    and the parameter gradients, gathered at q.
// States and arithmetic are FP32 (roift_relax_adjoint_dev) or FP64 (roift_relax_adjoint64_dev,
// the same kernels instantiated for double); the scalar gradients and the loss always
// accumulate in FP64.  The learnable scale map and the target are FP32 inputs in both.
#include <type_traits>

#include "ctx.hpp"

namespace roift {
namespace {

template <typename R>
struct DiffArgs {
    Dims d;
    R inv_s2;      // 1 / sigma^2
    R sigma;
    R alpha;
    R Dpq[27];
    FastDiv fx, fy;   // volumes below 2^31 voxels: coordinates by multiply-high division
    int fast;
};

// Coordinates of voxel p.
template <typename R>
__device__ inline void diff_coords(const DiffArgs<R> &a, long long p, int &x, int &y, int &z)
{
    if (a.fast) {
        const uint32_t t = a.fx.div((uint32_t)p);
        x = (int)((uint32_t)p - t * (uint32_t)a.d.xs);
        z = (int)a.fy.div(t);
        y = (int)(t - (uint32_t)z * (uint32_t)a.d.ys);
    } else {
        const long long xy = (long long)a.d.xs * a.d.ys;
        x = (int)(p % a.d.xs);
        y = (int)((p / a.d.xs) % a.d.ys);
        z = (int)(p / xy);
    }
}

// x / Dpq[k]: 6-adjacency has Dpq = 1, and dividing by 1 is exact, so it is skipped there.
template <int NADJ, typename R>
__device__ inline R over_dpq(R x, const DiffArgs<R> &a, int k)
{
    if constexpr (NADJ == 27) return x / a.Dpq[k];
    else return x;
}

template <typename R>
__device__ inline R rexp(R x)
{
    if constexpr (std::is_same<R, float>::value) return expf(x);
    else return exp(x);
}

template <int NADJ, typename IT, typename R>
__device__ inline R dweight(const IT *img, const float *scale, const DiffArgs<R> &a, int k, long long p, long long q,
                            R fq, R &g, R &sg, R &pi)
{
    const R d = (R)((int)img[p] - (int)img[q]);
    g = rexp<R>(-(d * d) * a.inv_s2);
    sg = d > (R)0 ? (R)1 : (d < (R)0 ? (R)-1 : (R)0);
    pi = (R)1 + a.alpha * sg * ((R)2 * fq - (R)1);
    return over_dpq<NADJ, R>((R)scale[q] * g * pi, a, k);
}

template <typename R>
__global__ void k_diff_init(const int32_t *__restrict__ label, long long n, int lab_obj, R *__restrict__ f,
                            uint8_t *__restrict__ seed)
{
    for (long long p = gtid(); p < n; p += gstride()) {
        f[p] = (label[p] == lab_obj) ? (R)0 : (R)1;
        seed[p] = 0;
    }
}

__global__ void k_diff_seeds(const int32_t *__restrict__ S, int ns, long long n, uint8_t *__restrict__ seed,
                             int *__restrict__ bad)
{
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int v = S[i];
        if (v >= 0 && v < n) seed[v] = 1;
        else atomicOr(bad, 1);
    }
}

template <int NADJ, typename IT, typename R>
__global__ __launch_bounds__(256) void k_diff_fwd(const IT *__restrict__ img, const float *__restrict__ scale,
                                                   DiffArgs<R> a, const uint8_t *__restrict__ seed,
                                                   const R *__restrict__ fin, R *__restrict__ fout)
{
    using A = Adjacency<NADJ>;
    const Dims d = a.d;
    const long long xy = (long long)d.xs * d.ys;
    for (long long p = gtid(); p < d.n; p += gstride()) {
        R v = fin[p];
        if (!seed[p]) {
            int x, y, z;
            diff_coords(a, p, x, y, z);
            R num = 0, den = 0;
#pragma unroll
            for (int k = 1; k < NADJ; k++) {
                const int qx = x + A::dx(k), qy = y + A::dy(k), qz = z + A::dz(k);
                if (!in_volume(d, qx, qy, qz)) continue;
                const long long q = qx + (long long)qy * d.xs + (long long)qz * xy;
                const R fq = fin[q];
                R g, sg, pi;
                const R w = dweight<NADJ, IT, R>(img, scale, a, k, p, q, fq, g, sg, pi);
                num += w * fq;
                den += w;
            }
            if (den != (R)0) v = num / den;
        }
        fout[p] = v;
    }
}

// Loss and its gradient with respect to the final f.
template <typename R>
__global__ __launch_bounds__(256) void k_diff_loss(const R *__restrict__ f, const float *__restrict__ target,
                                                    long long n, R eps, R *__restrict__ lam, double *__restrict__ part)
{
    double acc = 0.0;
    for (long long p = gtid(); p < n; p += gstride()) {
        R o = (R)1 - f[p];
        bool clamped = false;
        if (o < eps) { o = eps; clamped = true; }
        if (o > (R)1 - eps) { o = (R)1 - eps; clamped = true; }
        const float t = target[p];
        acc -= (double)t * log((double)o) + (1.0 - (double)t) * log(1.0 - (double)o);
        lam[p] = clamped ? (R)0 : (R)(((double)t / o - (1.0 - (double)t) / (1.0 - (double)o)) / (double)n);
    }
    acc = block_reduce(acc, OpAdd());
    if (threadIdx.x == 0) part[blockIdx.x] = acc / (double)n;   // summed by k_diff_sum
}

// out[0] += the sum of part[0 .. nb), in block order (one workgroup)
__global__ __launch_bounds__(256) void k_diff_sum(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
    s = block_reduce(s, OpAdd());
    if (threadIdx.x == 0) out[0] += s;
}

template <int NADJ, typename IT, typename R>
__global__ __launch_bounds__(256) void k_diff_bwd1(const IT *__restrict__ img, const float *__restrict__ scale,
                                                    DiffArgs<R> a, const uint8_t *__restrict__ seed,
                                                    const R *__restrict__ fin, const R *__restrict__ lam,
                                                    R *__restrict__ c, R *__restrict__ lam_out)
{
    using A = Adjacency<NADJ>;
    const Dims d = a.d;
    const long long xy = (long long)d.xs * d.ys;
    for (long long p = gtid(); p < d.n; p += gstride()) {
        R cp = 0, self = 0;
        const R l = lam[p];
        if (!seed[p] && l != (R)0) {
            int x, y, z;
            diff_coords(a, p, x, y, z);
            R den = 0;
#pragma unroll
            for (int k = 1; k < NADJ; k++) {
                const int qx = x + A::dx(k), qy = y + A::dy(k), qz = z + A::dz(k);
                if (!in_volume(d, qx, qy, qz)) continue;
                const long long q = qx + (long long)qy * d.xs + (long long)qz * xy;
                R g, sg, pi;
                den += dweight<NADJ, IT, R>(img, scale, a, k, p, q, fin[q], g, sg, pi);
            }
            if (den != (R)0) cp = l / den;
            else self = l;
        }
        c[p] = cp;
        lam_out[p] = self;
    }
}

template <int NADJ, typename IT, typename R>
__global__ __launch_bounds__(256) void k_diff_bwd2(const IT *__restrict__ img, const float *__restrict__ scale,
                                                    DiffArgs<R> a, const R *__restrict__ fin,
                                                    const R *__restrict__ fout, const R *__restrict__ c,
                                                    R *__restrict__ lam_out, R *__restrict__ gscale,
                                                    double *__restrict__ part)
{
    using A = Adjacency<NADJ>;
    const Dims d = a.d;
    const long long xy = (long long)d.xs * d.ys;
    double gsig = 0.0, galp = 0.0;
    const R s3 = a.sigma * a.sigma * a.sigma;
    const double k2s3 = 2.0 / (double)s3;
    for (long long q = gtid(); q < d.n; q += gstride()) {
        int x, y, z;
        diff_coords(a, q, x, y, z);
        const R fq = fin[q];
        R lq = lam_out[q], gs = 0;
#pragma unroll
        for (int k = 1; k < NADJ; k++) {
            const int px = x + A::dx(k), py = y + A::dy(k), pz = z + A::dz(k);
            if (!in_volume(d, px, py, pz)) continue;
            const long long p = px + (long long)py * d.xs + (long long)pz * xy;
            const R cp = c[p];
            if (cp == (R)0) continue;
            // the arc p -> q is p's neighbour in the opposite direction; Dpq is symmetric
            R g, sg, pi;
            const R w = dweight<NADJ, IT, R>(img, scale, a, k, p, q, fq, g, sg, pi);
            const R common = fq - fout[p];
            const R dwdf = over_dpq<NADJ, R>((R)scale[q] * g * a.alpha * sg * (R)2, a, k);
            lq += cp * (w + common * dwdf);
            gs += over_dpq<NADJ, R>(cp * common * g * pi, a, k);
            const R dd = (R)((int)img[p] - (int)img[q]);
            gsig += (double)(cp * common) * (double)w * (double)dd * (double)dd * k2s3;
            galp += (double)(cp * common) * (double)over_dpq<NADJ, R>((R)scale[q] * g * sg * ((R)2 * fq - (R)1), a, k);
        }
        lam_out[q] = lq;
        gscale[q] += gs;
    }
    // the block's sums join its own partial slots (the launches are ordered on one stream): one
    // device atomic per block on two shared addresses serialised and took most of the pass
    gsig = block_reduce(gsig, OpAdd());
    galp = block_reduce(galp, OpAdd());
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] += gsig;
        part[2 * blockIdx.x + 1] += galp;
    }
}

// gpar[0..1] += the sums of the nb blocks' partial slots, in block order (one workgroup)
__global__ __launch_bounds__(256) void k_diff_parts(const double *__restrict__ part, int nb, double *__restrict__ gpar)
{
    double s0 = 0.0, s1 = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        s0 += part[2 * i];
        s1 += part[2 * i + 1];
    }
    s0 = block_reduce(s0, OpAdd());
    s1 = block_reduce(s1, OpAdd());
    if (threadIdx.x == 0) {
        gpar[0] += s0;
        gpar[1] += s1;
    }
}

template <int NADJ, typename IT, typename R>
void diff_run(roift_ctx &c, const DiffArgs<R> &a, const void *d_image, const int32_t *d_seeds, int n_seeds,
              const int32_t *d_label, int lab_obj, int niter, int ckpt, const float *d_scale, const float *d_target,
              R eps, double *h_loss, R *d_gscale, double *h_gsig, double *h_galp, R *d_memb)
{
    hipStream_t st = c.stream;
    const long long n = a.d.n;
    const size_t rb = sizeof(R) * (size_t)n;
    const IT *img = (const IT *)d_image;
    const unsigned G = grid_for(n, 256, 8192);
    const int nseg = niter > 0 ? (niter + ckpt - 1) / ckpt : 0;
    uint8_t *seed = buf<uint8_t>(c, "diff.seed", n);
    R *ck = buf<R>(c, "diff.ckpt", (size_t)std::max(nseg, 1) * n);
    R *seg = buf<R>(c, "diff.seg", (size_t)(ckpt + 1) * n);
    R *lam = buf<R>(c, "diff.lam", n), *lam2 = buf<R>(c, "diff.lam2", n);
    R *cc = buf<R>(c, "diff.c", n);
    double *dacc = buf<double>(c, "diff.acc", 4);
    double *part = buf<double>(c, "diff.part", 2 * (size_t)G);
    fill<double>(c, part, 2 * (long long)G, 0.0);
    int *bad = reinterpret_cast<int *>(dacc + 3);
    const double zero[4] = {0, 0, 0, 0};
    ROIFT_HIP_CHECK(hipMemcpyAsync(dacc, zero, sizeof(zero), hipMemcpyHostToDevice, st));
    // forward, keeping the state at the start of every segment
    hipLaunchKernelGGL(k_diff_init<R>, dim3(G), dim3(256), 0, st, d_label, n, lab_obj, seg, seed);
    if (n_seeds > 0)
        hipLaunchKernelGGL(k_diff_seeds, dim3(grid_for(n_seeds)), dim3(256), 0, st, d_seeds, n_seeds, n, seed, bad);
    R *cur = seg, *nxt = seg + n;
    for (int it = 0; it < niter; it++) {
        if (it % ckpt == 0)
            ROIFT_HIP_CHECK(hipMemcpyAsync(ck + (size_t)(it / ckpt) * n, cur, rb, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL((k_diff_fwd<NADJ, IT, R>), dim3(G), dim3(256), 0, st, img, d_scale, a, seed, cur, nxt);
        std::swap(cur, nxt);
    }
    if (d_memb) ROIFT_HIP_CHECK(hipMemcpyAsync(d_memb, cur, rb, hipMemcpyDeviceToDevice, st));
    double *lpart = buf<double>(c, "diff.lpart", (size_t)G);
    hipLaunchKernelGGL(k_diff_loss<R>, dim3(G), dim3(256), 0, st, cur, d_target, n, eps, lam, lpart);
    hipLaunchKernelGGL(k_diff_sum, dim3(1), dim3(256), 0, st, lpart, (int)G, dacc);
    fill<R>(c, d_gscale, n, (R)0);
    ROIFT_HIP_CHECK(hipGetLastError());
    // backward, segment by segment
    for (int s = nseg - 1; s >= 0; s--) {
        const int k0 = s * ckpt, k1 = std::min(niter, k0 + ckpt), m = k1 - k0;
        ROIFT_HIP_CHECK(hipMemcpyAsync(seg, ck + (size_t)s * n, rb, hipMemcpyDeviceToDevice, st));
        for (int j = 0; j < m; j++)
            hipLaunchKernelGGL((k_diff_fwd<NADJ, IT, R>), dim3(G), dim3(256), 0, st, img, d_scale, a, seed,
                               seg + (size_t)j * n, seg + (size_t)(j + 1) * n);
        for (int j = m - 1; j >= 0; j--) {
            const R *fin = seg + (size_t)j * n, *fout = seg + (size_t)(j + 1) * n;
            hipLaunchKernelGGL((k_diff_bwd1<NADJ, IT, R>), dim3(G), dim3(256), 0, st, img, d_scale, a, seed, fin,
                               lam, cc, lam2);
            hipLaunchKernelGGL((k_diff_bwd2<NADJ, IT, R>), dim3(G), dim3(256), 0, st, img, d_scale, a, fin, fout,
                               cc, lam2, d_gscale, part);
            std::swap(lam, lam2);
        }
    }
    hipLaunchKernelGGL(k_diff_parts, dim3(1), dim3(256), 0, st, part, (int)G, dacc + 1);
    ROIFT_HIP_CHECK(hipGetLastError());
    double h[4];
    ROIFT_HIP_CHECK(hipMemcpyAsync(h, dacc, sizeof(h), hipMemcpyDeviceToHost, st));
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    ROIFT_REQUIRE(*reinterpret_cast<int *>(&h[3]) == 0, ROIFT_EINVAL, "seed index out of range");
    if (h_loss) *h_loss = h[0];
    if (h_gsig) *h_gsig = h[1];
    if (h_galp) *h_galp = h[2];
}

template <typename R>
void diff_dispatch(roift_ctx &c, int nadj, R sigma, R alpha, const Dims &d, const void *d_image,
                   int image_dtype, const int32_t *d_seeds, int n_seeds, const int32_t *d_label, int lab_obj,
                   int niter, int ckpt, const float *d_scale, const float *d_target, R eps, double *h_loss,
                   R *d_gscale, double *h_gsig, double *h_galp, R *d_memb)
{
    DiffArgs<R> a{};
    a.d = d;
    a.sigma = sigma;
    a.inv_s2 = (R)1 / (sigma * sigma);
    a.alpha = alpha;
    a.fast = d.n < (1ll << 31) ? 1 : 0;
    if (a.fast) {
        a.fx = FastDiv((uint32_t)d.xs);
        a.fy = FastDiv((uint32_t)d.ys);
    }
    for (int k = 0; k < 27; k++) a.Dpq[k] = 1;
    if (nadj == 27)
        for (int k = 1; k < 27; k++) {
            const int ax = Adjacency<27>::dx(k), ay = Adjacency<27>::dy(k), az = Adjacency<27>::dz(k);
            a.Dpq[k] = (R)std::sqrt((R)(ax * ax + ay * ay + az * az));
        }
    const bool i16 = image_dtype == ROIFT_I16;
    const R e = eps;
    if (nadj == 7) {
        if (i16) diff_run<7, int16_t, R>(c, a, d_image, d_seeds, n_seeds, d_label, lab_obj, niter, ckpt, d_scale,
                                         d_target, e, h_loss, d_gscale, h_gsig, h_galp, d_memb);
        else diff_run<7, int32_t, R>(c, a, d_image, d_seeds, n_seeds, d_label, lab_obj, niter, ckpt, d_scale,
                                     d_target, e, h_loss, d_gscale, h_gsig, h_galp, d_memb);
    } else {
        if (i16) diff_run<27, int16_t, R>(c, a, d_image, d_seeds, n_seeds, d_label, lab_obj, niter, ckpt, d_scale,
                                          d_target, e, h_loss, d_gscale, h_gsig, h_galp, d_memb);
        else diff_run<27, int32_t, R>(c, a, d_image, d_seeds, n_seeds, d_label, lab_obj, niter, ckpt, d_scale,
                                      d_target, e, h_loss, d_gscale, h_gsig, h_galp, d_memb);
    }
}

}  // namespace

int run_relax_diff(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype,
                   const int32_t *d_seeds, int n_seeds, const int32_t *d_label, int lab_obj, int niter, int ckpt,
                   const float *d_scale, double sigma, double alpha, const float *d_target, double eps,
                   double *h_loss, void *d_gscale, double *h_gsig, double *h_galp, void *d_memb, bool fp64)
{
    ROIFT_REQUIRE(nadj == 7 || nadj == 27, ROIFT_EUNSUPPORTED, "adjacency must be Spheric(1.0) or Spheric(1.75)");
    ROIFT_REQUIRE(niter >= 0 && ckpt >= 1 && n_seeds >= 0, ROIFT_EINVAL, "bad iteration / checkpoint / seed count");
    ROIFT_REQUIRE(sigma > 0.0 && eps > 0.0 && eps < 0.5, ROIFT_EINVAL, "sigma must be > 0 and eps in (0, 0.5)");
    ROIFT_REQUIRE(d_scale && d_target && d_gscale, ROIFT_EINVAL, "scale, target and gradient buffers are required");
    ev_record(c, 12);
    if (fp64)
        diff_dispatch<double>(c, nadj, sigma, alpha, d, d_image, image_dtype, d_seeds, n_seeds, d_label, lab_obj,
                              niter, ckpt, d_scale, d_target, eps, h_loss, (double *)d_gscale, h_gsig, h_galp,
                              (double *)d_memb);
    else   // the FP32 entry's float arguments, unchanged by the round trip through double
        diff_dispatch<float>(c, nadj, (float)sigma, (float)alpha, d, d_image, image_dtype, d_seeds, n_seeds, d_label,
                             lab_obj, niter, ckpt, d_scale, d_target, (float)eps, h_loss, (float *)d_gscale, h_gsig,
                             h_galp, (float *)d_memb);
    ev_record(c, 13);
    c.ktimes[2] = ev_ms(c, 12, 13);
    c.ktimes[3] = 3 * niter;
    return 0;
}

}  // namespace roift
