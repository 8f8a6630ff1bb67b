// blur.hip -- gft::Scene32::GaussianBlur (gft_filtering3.cpp:385-397) on the GPU.
//
// The reference convolves with Kernel3::SphericalGaussian(Spheric(1.0), s = 10, f = 1),
// normalised (gft_filtering3.cpp:72-126, :154-169): seven taps in Spheric order (centre,
// -z, -y, -x, +x, +y, +z), each voxel summing (float)I[q] * K[i] in that order over the
// taps inside the volume (no renormalisation at the border) and rounding with gft's ROUND
// (gft_common.h:120) (Convolution, gft_filtering3.cpp:321-346).  The CLI blurs twice
// (oiftrelax.cpp:368-375).
//
// Both passes run fused in one z-marching kernel: a workgroup owns a 64x16 column tile
// and a chunk of planes, keeps three input planes (2-voxel halo) and three first-pass
// planes (1-voxel halo) in LDS rings, and writes only the second pass, so the pair costs
// one HBM read and one write per voxel.  An odd pass count ends with the single-pass form
// of the same kernel.  This file is compiled without FMA contraction: every product and
// sum rounds as the reference's FP32 code does, so the result is bit-exact.
#include <algorithm>
#include <cmath>

#include "ctx.hpp"

namespace roift {
namespace {

constexpr int BX = 64, BY = 16;            // output tile (x, y)
constexpr int IX = BX + 4, IY = BY + 4;    // input plane with a 2-voxel halo
constexpr int PX = BX + 2, PY = BY + 2;    // first-pass plane with a 1-voxel halo
constexpr int NI = IX * IY, NP = PX * PY;
constexpr int THREADS = 256;               // 64 x 4
constexpr int NLOAD = (NI + THREADS - 1) / THREADS;

struct BlurArgs {
    int xs, ys, zs;
    unsigned ntx, nty, nzc;   // tiles in x and y, plane chunks in z
    int zc;                   // planes per chunk
    int xcd;                  // deal consecutive tiles to one XCD
    float k0, kz, ky, kx;     // normalised taps: centre, z, y and x neighbours
};

// ROUND(x) of gft_common.h:120 is (x < 0) ? (int)(x - 0.5) : (int)(x + 0.5) with a double
// 0.5, i.e. rounding half away from zero.  In FP32 that is trunc(x + copysign(h, x)) with
// h = 0.5 - 2^-25, the float just below 0.5: for x = n + f (n integral, 0 <= f < 1, x >= 0,
// spacing u of the floats at x), f >= 1/2 puts the exact sum within 2^-25 < u/2 of n + 1 or
// above it and below n + 3/2, so it rounds to [n + 1, n + 3/2); f <= 1/2 - u puts it below
// n + 1 - u, so it rounds to at most n + 1 - u; |x| < 1/2 gives at most 1 - 2^-24 and
// |x| = 1/2 ties to 1; |x| >= 2^23 is integral and h < u/2 leaves it unchanged.  Negative x
// mirrors this.  Three instructions instead of the six of roundf.
__device__ __forceinline__ float gft_round(float conv)
{
    return __builtin_truncf(conv + __builtin_copysignf(0x1.fffffep-2f, conv));
}

// One voxel: FP32 products and sums in the reference's tap order (centre, -z, -y, -x, +x,
// +y, +z).  Values outside the volume are staged as +0.0f, and (float)q * k with q = +0.0f
// adds +0.0f, which leaves the running sum unchanged (it starts at +0.0f and is never
// -0.0f), exactly as skipping the tap does.  The taps are positive.
__device__ __forceinline__ float blur_voxel(const BlurArgs &a, float c, float zm, float ym, float xm, float xp,
                                            float yp, float zp)
{
    // the reference starts from conv = 0.0f; 0.0f + c * k0 == c * k0 because c, an
    // integer converted to float, is never -0.0f
    float conv = c * a.k0;
    conv += zm * a.kz;
    conv += ym * a.ky;
    conv += xm * a.kx;
    conv += xp * a.kx;
    conv += yp * a.ky;
    conv += zp * a.kz;
    return gft_round(conv);
}

__device__ __forceinline__ int ring(int k) { return (k + 3) % 3; }   // k >= -2

template <typename IT, bool TWO>
__global__ __launch_bounds__(THREADS) void k_blur(const IT *__restrict__ in, IT *__restrict__ out, BlurArgs a)
{
    // planes as floats: (float)I[q] is formed once, when a plane is staged; first-pass
    // results are integral floats (the reference converts its int result back to float)
    __shared__ float s_in[3][NI];
    __shared__ float s_p1[TWO ? 3 : 1][TWO ? NP : 1];
    const unsigned ntiles = a.ntx * a.nty * a.nzc;
    unsigned tile = blockIdx.x;
    if (a.xcd) tile = (blockIdx.x % 8u) * ((ntiles + 7u) / 8u) + blockIdx.x / 8u;
    if (tile >= ntiles) return;   // uniform per workgroup: no barrier is skipped by part of it
    const int x0 = (int)(tile % a.ntx) * BX, y0 = (int)((tile / a.ntx) % a.nty) * BY;
    const int z0 = (int)(tile / (a.ntx * a.nty)) * a.zc, z1 = min(a.zs, z0 + a.zc);
    const int tid = threadIdx.x;
    const size_t xy = (size_t)a.xs * (size_t)a.ys;

    // Input planes (2-voxel halo) are fetched two planes ahead of use into two register
    // sets; the z loop is unrolled by two so each set keeps a fixed role per step.  A
    // thread's in-plane offsets and their validity do not depend on the plane: they are
    // formed once, addresses are clamped into the volume and every load is issued without
    // a branch (so the loads of a plane are all in flight together), and values outside
    // the volume are replaced by 0 when the plane is staged.
    uint32_t off[NLOAD];
    uint32_t xymask = 0;
#pragma unroll
    for (int r = 0; r < NLOAD; r++) {
        const int idx = tid + r * THREADS;
        const int gx = x0 - 2 + idx % IX, gy = y0 - 2 + idx / IX;
        const bool valid = idx < NI && gx >= 0 && gx < a.xs && gy >= 0 && gy < a.ys;
        off[r] = (uint32_t)min(max(gy, 0), a.ys - 1) * (uint32_t)a.xs + (uint32_t)min(max(gx, 0), a.xs - 1);
        xymask |= (uint32_t)valid << r;
    }
    int32_t ra[NLOAD], rb[NLOAD];
    uint32_t ma = 0, mb = 0;
    auto fetch = [&](int k, int32_t (&reg)[NLOAD], uint32_t &mask) {
        const IT *plane = in + (size_t)min(max(k, 0), a.zs - 1) * xy;
#pragma unroll
        for (int r = 0; r < NLOAD; r++) reg[r] = (int32_t)plane[off[r]];
        mask = (k >= 0 && k < a.zs) ? xymask : 0u;
    };
    auto stash = [&](int k, const int32_t (&reg)[NLOAD], uint32_t mask) {
        float *s = s_in[ring(k)];
#pragma unroll
        for (int r = 0; r < NLOAD; r++) {
            const int idx = tid + r * THREADS;
            if (idx < NI) s[idx] = ((mask >> r) & 1u) ? (float)reg[r] : 0.0f;
        }
    };
    // Output positions of this thread: (tx, ty0 + r * ROWS) for r < BY / ROWS, with their
    // in-plane offsets in HBM and whether they lie inside the volume.
    constexpr int ROWS = THREADS / BX, NOUT = BY / ROWS;
    const int tx = tid % BX, ty0 = tid / BX;
    const uint32_t obase = (uint32_t)min(y0 + ty0, a.ys - 1) * (uint32_t)a.xs + (uint32_t)min(x0 + tx, a.xs - 1);
    uint32_t omask = 0;
#pragma unroll
    for (int r = 0; r < NOUT; r++)
        omask |= (uint32_t)(x0 + tx < a.xs && y0 + ty0 + r * ROWS < a.ys) << r;
    // Output plane k from three planes (k-1, k, k+1) of a buffer with row pitch W whose
    // output tile starts at row/column H: the single pass reads the input ring (W = IX,
    // H = 2), the fused pair's second pass the first-pass ring (W = PX, H = 1).
    auto emit = [&](int k, const float *pm, const float *pc, const float *pp, int W, int H) {
        IT *o = out + (size_t)k * xy + obase;
#pragma unroll
        for (int r = 0; r < NOUT; r++) {
            const int s = (ty0 + r * ROWS + H) * W + tx + H;
            const float v = blur_voxel(a, pc[s], pm[s], pc[s - W], pc[s - 1], pc[s + 1], pc[s + W], pp[s]);
            if ((omask >> r) & 1u) o[(uint32_t)(r * ROWS) * (uint32_t)a.xs] = (IT)(int32_t)v;
        }
    };

    if constexpr (TWO) {
        // First-pass positions of this thread: region entries tid + r * THREADS of the
        // (PX x PY) plane, their centre offsets in the input ring and whether they lie in
        // the volume (entries outside it are staged as 0, so the second pass adds +0.0f).
        constexpr int NP1 = (NP + THREADS - 1) / THREADS;
        int p1s[NP1];
        uint32_t p1mask = 0;
#pragma unroll
        for (int r = 0; r < NP1; r++) {
            const int idx = min(tid + r * THREADS, NP - 1);
            const int i = idx % PX, j = idx / PX, gx = x0 - 1 + i, gy = y0 - 1 + j;
            p1s[r] = (j + 1) * IX + i + 1;
            p1mask |= (uint32_t)(gx >= 0 && gx < a.xs && gy >= 0 && gy < a.ys) << r;
        }
        auto pass1 = [&](int k) {   // first pass of plane k into the first-pass ring
            const float *cm = s_in[ring(k - 1)], *cc = s_in[ring(k)], *cp = s_in[ring(k + 1)];
            float *dst = s_p1[ring(k)];
            const uint32_t m = (k >= 0 && k < a.zs) ? p1mask : 0u;
#pragma unroll
            for (int r = 0; r < NP1; r++) {
                const int idx = tid + r * THREADS;
                if (r < NP1 - 1 || idx < NP) {
                    const int s = p1s[r];
                    const float v =
                        blur_voxel(a, cc[s], cm[s], cc[s - IX], cc[s - 1], cc[s + 1], cc[s + IX], cp[s]);
                    dst[idx] = ((m >> r) & 1u) ? v : 0.0f;
                }
            }
        };
        auto pass2 = [&](int k) { emit(k, s_p1[ring(k - 1)], s_p1[ring(k)], s_p1[ring(k + 1)], PX, 1); };
        // prologue: input planes z0-2 .. z0+1, first-pass planes z0-1 and z0
        fetch(z0 - 2, ra, ma); stash(z0 - 2, ra, ma);
        fetch(z0 - 1, ra, ma); stash(z0 - 1, ra, ma);
        fetch(z0, ra, ma); stash(z0, ra, ma);
        fetch(z0 + 1, ra, ma);
        fetch(z0 + 2, rb, mb);
        __syncthreads();
        pass1(z0 - 1);
        __syncthreads();
        stash(z0 + 1, ra, ma);   // overwrites plane z0-2, last read just above
        fetch(z0 + 3, ra, ma);
        __syncthreads();
        pass1(z0);
        __syncthreads();
        // step z: stash plane z+2 (held in `cur`), refill `cur` with plane z+4.  Input plane
        // z-1 (read by the first pass of plane z) and first-pass plane z-2 (read by the second
        // pass of plane z-1) were both released by a barrier.
        auto step = [&](int z, int32_t (&cur)[NLOAD], uint32_t &m) {
            stash(z + 2, cur, m);
            fetch(z + 4, cur, m);
            __syncthreads();
            pass1(z + 1);
            __syncthreads();
            pass2(z);
        };
        for (int z = z0; z < z1; z += 2) {   // rb holds plane z+2, ra plane z+3
            step(z, rb, mb);
            if (z + 1 < z1) step(z + 1, ra, ma);   // uniform per workgroup
        }
    } else {
        fetch(z0 - 1, ra, ma); stash(z0 - 1, ra, ma);
        fetch(z0, ra, ma); stash(z0, ra, ma);
        fetch(z0 + 1, ra, ma);
        fetch(z0 + 2, rb, mb);
        auto step = [&](int z, int32_t (&cur)[NLOAD], uint32_t &m) {
            stash(z + 1, cur, m);   // overwrites plane z-2, released by the barrier ending the last step
            fetch(z + 3, cur, m);
            __syncthreads();
            emit(z, s_in[ring(z - 1)], s_in[ring(z)], s_in[ring(z + 1)], IX, 2);
            __syncthreads();
        };
        for (int z = z0; z < z1; z += 2) {   // ra holds plane z+1, rb plane z+2
            step(z, ra, ma);
            if (z + 1 < z1) step(z + 1, rb, mb);
        }
    }
}

// Kernel3::SphericalGaussian(scn, R = 1, s = 10, f = 1) + Normalize over Spheric(1.0),
// computed on the host exactly as the reference's host code does (gft_filtering3.cpp:106-126,
// 72-83): displacements scaled by spacing / min spacing in float, exp of a float, float
// sum in tap order, float division.
void gauss_taps(const Dims &d, BlurArgs &a)
{
    const float R = 1.0f, s = 10.0f, f = 1.0f, R2 = R * R;
    const float dmin = std::min(d.sx, std::min(d.sy, d.sz));
    float val[7];
    for (int i = 0; i < 7; i++) {
        const float d_x = ((float)Adjacency<7>::dx(i) * d.sx) / dmin;
        const float d_y = ((float)Adjacency<7>::dy(i) * d.sy) / dmin;
        const float d_z = ((float)Adjacency<7>::dz(i) * d.sz) / dmin;
        const float r2 = d_x * d_x + d_y * d_y + d_z * d_z;
        val[i] = s * std::exp(-f * (r2 / R2));   // the float overload, as the reference's exp(float)
    }
    float wt = 0.0f;
    for (int i = 0; i < 7; i++) wt += val[i];
    a.k0 = val[0] / wt;
    a.kz = val[1] / wt;   // -z (the +z tap, i = 6, has the same value)
    a.ky = val[2] / wt;
    a.kx = val[3] / wt;
}

template <typename IT>
void launch_blur(roift_ctx &c, const Dims &d, const void *src, void *dst, bool two, const BlurArgs &a)
{
    const unsigned ntiles = a.ntx * a.nty * a.nzc;
    const dim3 grid(a.xcd ? 8u * ((ntiles + 7u) / 8u) : ntiles);
    if (two)
        hipLaunchKernelGGL((k_blur<IT, true>), grid, dim3(THREADS), 0, c.stream, (const IT *)src, (IT *)dst, a);
    else
        hipLaunchKernelGGL((k_blur<IT, false>), grid, dim3(THREADS), 0, c.stream, (const IT *)src, (IT *)dst, a);
    ROIFT_HIP_CHECK(hipGetLastError());
}

}  // namespace

int run_gaussian_blur(roift_ctx &c, const Dims &d, const void *d_in, int dtype, void *d_out, int passes)
{
    ROIFT_REQUIRE(dtype == ROIFT_I16 || dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16 or I32");
    ROIFT_REQUIRE(passes >= 0, ROIFT_EINVAL, "passes must be >= 0");
    ROIFT_REQUIRE(d_in && d_out, ROIFT_EINVAL, "null image pointer");
    const size_t esz = dtype == ROIFT_I16 ? 2 : 4;
    const size_t bytes = (size_t)d.n * esz;
    ev_record(c, 14);
    if (passes == 0) {
        if (d_in != d_out) ROIFT_HIP_CHECK(hipMemcpyAsync(d_out, d_in, bytes, hipMemcpyDeviceToDevice, c.stream));
        ev_record(c, 15);
        c.ktimes[4] = ev_ms(c, 14, 15);
        return ROIFT_OK;
    }
    BlurArgs a{};
    a.xs = d.xs;
    a.ys = d.ys;
    a.zs = d.zs;
    gauss_taps(d, a);
    a.ntx = (unsigned)((d.xs + BX - 1) / BX);
    a.nty = (unsigned)((d.ys + BY - 1) / BY);
    // planes per workgroup: ROIFT_PARAM_BLUR_PLANES, else enough chunks for ~1024 workgroups
    int zc = c.tune_blur_planes;
    if (zc <= 0) {
        const long long cols = (long long)a.ntx * a.nty;
        zc = (int)std::max(4LL, std::min<long long>(d.zs, (d.zs * cols + 1023) / 1024));
    }
    a.zc = std::min(zc, d.zs);
    a.nzc = (unsigned)((d.zs + a.zc - 1) / a.zc);
    a.xcd = 1;
    ROIFT_REQUIRE((unsigned long long)a.ntx * a.nty * a.nzc + 8 < (1ull << 31), ROIFT_EUNSUPPORTED,
                  "volume too large for the blur grid");

    // Launch sequence: passes / 2 fused pairs, then one single pass if odd.  The last
    // launch writes d_out; earlier ones alternate with a scratch volume.  No launch may
    // read and write the same volume, so an in-place call first copies its input aside.
    const int nl = passes / 2 + (passes & 1);
    const void *src = d_in;
    if (d_in == d_out) {
        void *cp = buf<char>(c, "blur.src", bytes);
        ROIFT_HIP_CHECK(hipMemcpyAsync(cp, d_in, bytes, hipMemcpyDeviceToDevice, c.stream));
        src = cp;
    }
    void *tmp = nl > 1 ? buf<char>(c, "blur.tmp", bytes) : nullptr;
    for (int l = 0; l < nl; l++) {
        void *dst = ((nl - 1 - l) % 2 == 0) ? d_out : tmp;
        const bool two = (l < passes / 2);
        if (dtype == ROIFT_I16) launch_blur<int16_t>(c, d, src, dst, two, a);
        else launch_blur<int32_t>(c, d, src, dst, two, a);
        src = dst;
    }
    ev_record(c, 15);
    c.ktimes[4] = ev_ms(c, 14, 15);
    return ROIFT_OK;
}

}  // namespace roift
