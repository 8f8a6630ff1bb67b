// stages.hip -- the CLI's stages on either side of the hot path, on the GPU, bit-exact:
//
//   Sobel gradient source    gft::Scene32::SobelFilter (gft_filtering3.cpp:509-552) and the
//                            experiments' normalisation to the image maximum
//                            (exp_gradient_weight.cpp:593-607)
//   object-seed distance     compute_bfs_edt (exp_geodesic_seeds.cpp:165-215)
//   boundary seeds           inject_boundary_seeds (oiftrelax.cpp:175-222) and its geodesic
//                            form (exp_geodesic_seeds.cpp:221-283)
//   conditional dilation     dilation_conditional (oiftrelax.cpp:87-169)
//
// Compiled without FMA contraction (the Sobel magnitude follows the reference's FP32 order).
#include <climits>
#include <cmath>

#include "compact.hpp"

namespace roift {

// k_cp_scan of compact.hpp: one workgroup, each thread scans a contiguous run of counts.
__global__ __launch_bounds__(1024) void k_cp_scan(uint32_t *__restrict__ blk, int nb)
{
    const int per = (nb + 1023) / 1024, t = (int)threadIdx.x;
    uint32_t s = 0;
    for (int i = 0; i < per; i++) {
        const int k = t * per + i;
        if (k < nb) s += blk[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<1024>(s, total);
    for (int i = 0; i < per; i++) {
        const int k = t * per + i;
        if (k < nb) {
            const uint32_t v = blk[k];
            blk[k] = run;
            run += v;
        }
    }
    if (t == 0) blk[nb] = total;
}

namespace {

// ---------------------------------------------------------------------------------
// Sobel
// ---------------------------------------------------------------------------------

__device__ __forceinline__ float next_up(float x) { return __uint_as_float(__float_as_uint(x) + 1u); }     // x >= 0
__device__ __forceinline__ float next_down(float x) { return x > 0.0f ? __uint_as_float(__float_as_uint(x) - 1u) : x; }

// RN(x / 3) in FP32 for x >= 0, as the reference's (double)x / 3.0 rounded to float (double
// rounding is innocuous for division: 53 >= 2*24 + 2).  The first estimate is within one
// float ulp; it is corrected by comparing x with 3 * (the midpoints to its neighbours),
// products that are exact in double (a midpoint has 25 significant bits).  No midpoint
// times 3 is a float, so there are no ties.
__device__ __forceinline__ float div3_rn(float x)
{
    float q = (float)((double)x / 3.0);
    const float up = next_up(q), dn = next_down(q);
    const double xd = (double)x;
    if (xd > 3.0 * (0.5 * ((double)q + (double)up))) q = up;
    else if (q > 0.0f && xd < 3.0 * (0.5 * ((double)q + (double)dn))) q = dn;
    return q;
}

// RN(sqrt(y)) in FP32 for y >= 0 (sqrtf), corrected the same way: the square of a midpoint
// has at most 50 significant bits, exact in double, and is never a float.
__device__ __forceinline__ float sqrt_rn(float y)
{
    float r = (float)__builtin_sqrt((double)y);
    const float up = next_up(r), dn = next_down(r);
    const double yd = (double)y, mu = 0.5 * ((double)r + (double)up), md = 0.5 * ((double)r + (double)dn);
    if (yd > mu * mu) r = up;
    else if (r > 0.0f && yd < md * md) r = dn;
    return r;
}

// ROUND(sqrtf((f0*f0 + f1*f1 + f2*f2) / 3.0)) of gft_filtering3.cpp:546-549.
__device__ __forceinline__ int32_t sobel_magnitude(int a0, int a1, int a2)
{
    const float f0 = (float)a0, f1 = (float)a1, f2 = (float)a2;
    const float sum = f0 * f0 + f1 * f1 + f2 * f2;
    const float g = sqrt_rn(div3_rn(sum));
    return (int32_t)((double)g + 0.5);   // gft's ROUND of a non-negative value
}

// Each thread computes SV consecutive linear indices: the 3 x 3 rows (offsets -xy, 0, +xy
// by -xs, 0, +xs) around them, each SV + 2 values wide, are loaded once and shared (9
// (SV + 2) loads instead of 54 SV).  Everything is in linear offsets, so the neighbours
// wrap across rows and planes exactly as the reference's addresses do.  The gradient is
// computed over the reference's range [di, N - di), di = 1 + xs + xs*ys, and is 0
// elsewhere (gft scenes are created zeroed).  maxes[2b] / [2b + 1] receive workgroup b's
// max G / max image.
constexpr int SV = 4;

template <typename IT>
__global__ __launch_bounds__(256) void k_sobel(const IT *__restrict__ img, Dims d, int32_t *__restrict__ G,
                                               int *__restrict__ maxes)
{
    const long long N = d.n, dy = d.xs, dz = (long long)d.xs * d.ys, di = 1 + dy + dz;
    int gmax = INT_MIN, imax = INT_MIN;
    const long long groups = (N + SV - 1) / SV;
    for (long long t = gtid(); t < groups; t += gstride()) {
        const long long i0 = t * SV;
        int32_t g[SV];
#pragma unroll
        for (int k = 0; k < SV; k++) {
            g[k] = 0;
            if (i0 + k < N) imax = max(imax, (int)img[i0 + k]);
        }
        if (i0 + SV - 1 >= di && i0 < N - di) {
            // v[r][c] = I[i0 - 1 + c + oz(r) + oy(r)], r = (z+1)*3 + (y+1); clamped addresses
            // only where an output lies outside [di, N - di) (its value is not used)
            int v[9][SV + 2];
#pragma unroll
            for (int r = 0; r < 9; r++) {
                const long long o = (long long)(r / 3 - 1) * dz + (long long)(r % 3 - 1) * dy;
#pragma unroll
                for (int c = 0; c < SV + 2; c++) {
                    const long long q = min(max(i0 - 1 + c + o, 0LL), N - 1);
                    v[r][c] = (int)img[q];
                }
            }
#pragma unroll
            for (int k = 0; k < SV; k++) {
                const long long i = i0 + k;
                if (i < di || i >= N - di) continue;
                // row r = (zz+1)*3 + (yy+1), column c = k + 1 + xx
                auto V = [&](int xx, int yy, int zz) { return v[(zz + 1) * 3 + (yy + 1)][k + 1 + xx]; };
                int a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
                for (int m = 0; m < 9; m++) {
                    const int n0 = m % 3 - 1, n1 = m / 3 - 1;
                    const int we = (n0 == 0 ? 2 : 1) * (n1 == 0 ? 2 : 1);   // 1 2 1 / 2 4 2 / 1 2 1
                    a0 += we * (V(n0, n1, 1) - V(n0, n1, -1));    // xy plane: z difference
                    a1 += we * (V(1, n1, n0) - V(-1, n1, n0));    // yz plane: x difference
                    a2 += we * (V(n0, 1, n1) - V(n0, -1, n1));    // xz plane: y difference
                }
                g[k] = sobel_magnitude(a0 >> 4, a1 >> 4, a2 >> 4);
            }
        }
#pragma unroll
        for (int k = 0; k < SV; k++)
            if (i0 + k < N) {
                G[i0 + k] = g[k];
                gmax = max(gmax, (int)g[k]);
            }
    }
    // workgroup maxima to this workgroup's own slots (no contended atomics); k_max_fold
    // folds them
    __shared__ int sg[4], si[4];
    for (int o = 32; o > 0; o >>= 1) {
        gmax = max(gmax, __shfl_xor(gmax, o));
        imax = max(imax, __shfl_xor(imax, o));
    }
    if (__lane_id() == 0) {
        sg[threadIdx.x / 64] = gmax;
        si[threadIdx.x / 64] = imax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        maxes[2 * blockIdx.x] = max(max(sg[0], sg[1]), max(sg[2], sg[3]));
        maxes[2 * blockIdx.x + 1] = max(max(si[0], si[1]), max(si[2], si[3]));
    }
}

// maxes[0..1] = elementwise maxima of the nb pairs part[2b], part[2b + 1] (one workgroup).
__global__ __launch_bounds__(1024) void k_max_fold(const int *__restrict__ part, int nb, int *__restrict__ maxes)
{
    int a = INT_MIN, b = INT_MIN;
    for (int i = threadIdx.x; i < nb; i += 1024) {
        a = max(a, part[2 * i]);
        b = max(b, part[2 * i + 1]);
    }
    __shared__ int sa[16], sb[16];
    for (int o = 32; o > 0; o >>= 1) {
        a = max(a, __shfl_xor(a, o));
        b = max(b, __shfl_xor(b, o));
    }
    if (__lane_id() == 0) {
        sa[threadIdx.x / 64] = a;
        sb[threadIdx.x / 64] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; w++) {
            a = max(a, sa[w]);
            b = max(b, sb[w]);
        }
        maxes[0] = a;
        maxes[1] = b;
    }
}

// G = (int)((long long)G * Wmax / Gmax) when both maxima are positive.
__global__ __launch_bounds__(256) void k_grad_normalize(int32_t *__restrict__ G, long long n,
                                                        const int *__restrict__ maxes)
{
    const long long gmax = maxes[0], wmax = maxes[1];
    if (gmax <= 0 || wmax <= 0) return;
    for (long long i = gtid(); i < n; i += gstride()) G[i] = (int32_t)((long long)G[i] * wmax / gmax);
}

// ---------------------------------------------------------------------------------
// Distance from the object seeds: on the unobstructed 6-connected grid the BFS distance
// is the L1 distance to the nearest source, computed exactly by three separable min-plus
// passes (a forward and a backward sweep per line).
// ---------------------------------------------------------------------------------

constexpr int32_t EDT_INF = 0x3FFFFFFF;

__global__ __launch_bounds__(256) void k_edt_sources(const int32_t *__restrict__ S, int nS,
                                                     const int32_t *__restrict__ label, long long n,
                                                     int32_t *__restrict__ dist)
{
    for (long long i = gtid(); i < nS; i += gstride()) {
        const long long p = S[i];
        if (p >= 0 && p < n && label[p] > 0) dist[p] = 0;
    }
}

// Lines along one axis: `lines` lines of length L, element k of line l at
// base(l) + k * step.  The last pass maps distances >= EDT_INF to INT_MAX (unreached).
template <int AXIS>
__global__ __launch_bounds__(256) void k_edt_pass(int32_t *__restrict__ dist, Dims d, bool last)
{
    const long long xy = (long long)d.xs * d.ys;
    const long long lines = AXIS == 0 ? (long long)d.ys * d.zs : (AXIS == 1 ? (long long)d.xs * d.zs : xy);
    const int L = AXIS == 0 ? d.xs : (AXIS == 1 ? d.ys : d.zs);
    const long long step = AXIS == 0 ? 1 : (AXIS == 1 ? d.xs : xy);
    for (long long l = gtid(); l < lines; l += gstride()) {
        const long long base = AXIS == 0 ? l * d.xs : (AXIS == 1 ? (l % d.xs) + (l / d.xs) * xy : l);
        int32_t run = EDT_INF;
        for (int k = 0; k < L; k++) {
            const long long p = base + k * step;
            run = min(min(dist[p], run + 1), EDT_INF);
            dist[p] = run;
        }
        run = EDT_INF;
        for (int k = L - 1; k >= 0; k--) {
            const long long p = base + k * step;
            run = min(min(dist[p], run + 1), EDT_INF);
            dist[p] = (last && run >= EDT_INF) ? INT_MAX : run;
        }
    }
}

// The x pass: one wavefront per row, as min-plus scans over 64-voxel chunks.  Forward,
// F(x) = x + min_{x' <= x} (d(x') - x'); backward on F, D(x) = -x + min_{x' >= x} (F(x') + x'),
// which is min_{x'} d(x') + |x - x'| -- the same values as the sequential sweeps, clamped
// at EDT_INF the same way.
__global__ __launch_bounds__(256) void k_edt_rows(int32_t *__restrict__ dist, Dims d)
{
    const long long rows = (long long)d.ys * d.zs;
    const int lane = __lane_id();
    constexpr int BIG = 0x7FFFFFFF - (1 << 16);   // above every d - x and F + x
    for (long long r = gtid() >> 6; r < rows; r += gstride() >> 6) {
        int32_t *row = dist + r * d.xs;
        int carry = BIG;
        for (int x0 = 0; x0 < d.xs; x0 += 64) {
            const int x = x0 + lane;
            int v = x < d.xs ? row[x] - x : BIG;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(v, o);
                if (lane >= o) v = min(v, t);
            }
            v = min(v, carry);
            carry = __shfl(v, 63);
            if (x < d.xs) row[x] = min(x + v, EDT_INF);
        }
        carry = BIG;
        for (int x0 = ((d.xs - 1) / 64) * 64; x0 >= 0; x0 -= 64) {
            const int x = x0 + lane;
            int v = x < d.xs ? row[x] + x : BIG;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_down(v, o);
                if (lane + o < 64) v = min(v, t);
            }
            v = min(v, carry);
            carry = __shfl(v, 0);
            if (x < d.xs) row[x] = min(v - x, EDT_INF);
        }
    }
}

// ---------------------------------------------------------------------------------
// Boundary seeds
// ---------------------------------------------------------------------------------

struct VoxelXYZ {
    int xs, ys, zs;
    __device__ __forceinline__ void of(long long p, int &x, int &y, int &z) const
    {
        const uint32_t u = (uint32_t)p, xy = (uint32_t)xs * (uint32_t)ys;
        z = (int)(u / xy);
        const uint32_t r = u - (uint32_t)z * xy;
        y = (int)(r / (uint32_t)xs);
        x = (int)(r - (uint32_t)y * (uint32_t)xs);
    }
    __device__ __forceinline__ bool on_face(int x, int y, int z) const
    {
        return x == 0 || x == xs - 1 || y == 0 || y == ys - 1 || z == 0 || z == zs - 1;
    }
};

// inject_boundary_seeds: grid voxels (x, y, z all multiples of stride) on a face, unlabelled.
struct PredStride {
    VoxelXYZ v;
    const int32_t *label;
    int stride;
    __device__ bool operator()(long long p) const
    {
        int x, y, z;
        v.of(p, x, y, z);
        if (x % stride || y % stride || z % stride || !v.on_face(x, y, z)) return false;
        return label[p] == NIL;
    }
};

// inject_boundary_seeds_geodesic: face voxels, unlabelled, farther than 2 stride from the
// object seeds, on the effective-stride grid of any one axis.
struct PredGeodesic {
    VoxelXYZ v;
    const int32_t *label, *dist;
    int stride;
    __device__ bool operator()(long long p) const
    {
        int x, y, z;
        v.of(p, x, y, z);
        if (!v.on_face(x, y, z) || label[p] != NIL) return false;
        const int dd = dist[p];
        if (dd <= stride * 2) return false;
        const int es = dd <= stride ? max(1, stride / 2) : (dd <= 3 * stride ? stride : stride * 2);
        return x % es == 0 || y % es == 0 || z % es == 0;
    }
};

struct EmitSeed {
    int32_t *label, *out;
    __device__ void operator()(long long p, uint32_t r) const
    {
        label[p] = 0;
        out[r] = (int32_t)p;
    }
};

// ---------------------------------------------------------------------------------
// Conditional dilation
// ---------------------------------------------------------------------------------

constexpr int DIL_MAXN = 125;
struct DilAdj {
    int n;
    int dx[DIL_MAXN], dy[DIL_MAXN], dz[DIL_MAXN];
};

// border[p] = label 0 with a Spheric(r) neighbour of label 1 (the reference marks the
// label-0 neighbours of every label-1 voxel; the ball is symmetric, so this is the same set).
__global__ __launch_bounds__(256) void k_dil_border(const int32_t *__restrict__ label, Dims d, DilAdj A,
                                                    uint8_t *__restrict__ border)
{
    const VoxelXYZ v{d.xs, d.ys, d.zs};
    for (long long p = gtid(); p < d.n; p += gstride()) {
        uint8_t b = 0;
        if (label[p] == 0) {
            int x, y, z;
            v.of(p, x, y, z);
            for (int i = 1; i < A.n && !b; i++) {
                const int qx = x + A.dx[i], qy = y + A.dy[i], qz = z + A.dz[i];
                if (qx < 0 || qx >= d.xs || qy < 0 || qy >= d.ys || qz < 0 || qz >= d.zs) continue;
                b = label[qx + (long long)qy * d.xs + (long long)qz * d.xs * d.ys] == 1;
            }
        }
        border[p] = b;
    }
}

// Spheric(1.0) border, z-marching: a workgroup owns a 64 x 4 column tile and a chunk of
// planes, each thread keeps its column's labels at z-1, z, z+1 in registers and reads the
// four in-plane neighbours (rows the workgroup shares in cache).
constexpr int DB_X = 64, DB_Y = 4, DB_Z = 32;
__global__ __launch_bounds__(DB_X *DB_Y) void k_dil_border6(const int32_t *__restrict__ label, Dims d,
                                                          uint8_t *__restrict__ border)
{
    const int x = blockIdx.x * DB_X + threadIdx.x, y = blockIdx.y * DB_Y + threadIdx.y;
    const int z0 = blockIdx.z * DB_Z, z1 = min(d.zs, z0 + DB_Z);
    if (x >= d.xs || y >= d.ys) return;   // no workgroup barrier below
    const size_t xy = (size_t)d.xs * d.ys, col = (size_t)y * d.xs + x;
    int lm = z0 > 0 ? label[(size_t)(z0 - 1) * xy + col] : -1;
    int lc = label[(size_t)z0 * xy + col];
    for (int z = z0; z < z1; z++) {
        const size_t p = (size_t)z * xy + col;
        const int lp = z + 1 < d.zs ? label[p + xy] : -1;
        bool b = false;
        if (lc == 0)
            b = lm == 1 || lp == 1 || (x > 0 && label[p - 1] == 1) || (x + 1 < d.xs && label[p + 1] == 1) ||
                (y > 0 && label[p - d.xs] == 1) || (y + 1 < d.ys && label[p + d.xs] == 1);
        border[p] = b;
        lm = lc;
        lc = lp;
    }
}

struct PredFlag {
    const uint8_t *flag;
    __device__ bool operator()(long long p) const { return flag[p] != 0; }
};

// the border list in raster order: position and order-preserving key of the image value
template <typename IT>
struct EmitBorder {
    const IT *img;
    int32_t *pos;
    uint32_t *key;
    __device__ void operator()(long long p, uint32_t r) const
    {
        pos[r] = (int32_t)p;
        key[r] = (uint32_t)(int32_t)img[p] ^ 0x80000000u;
    }
};

// 256-bin histogram of key digit (key >> shift) & 255 over the keys matching prefix/mask.
__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t *__restrict__ key, long long n, uint32_t prefix,
                                                    uint32_t mask, int shift, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (long long i = gtid(); i < n; i += gstride()) {
        const uint32_t k = key[i];
        if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_dil_below(const uint32_t *__restrict__ key, const int32_t *__restrict__ pos,
                                                   long long n, uint32_t t, int32_t *__restrict__ label)
{
    for (long long i = gtid(); i < n; i += gstride())
        if (key[i] < t) label[pos[i]] = 1;
}

__global__ __launch_bounds__(256) void k_dil_all(const int32_t *__restrict__ pos, long long n,
                                                 int32_t *__restrict__ label)
{
    for (long long i = gtid(); i < n; i += gstride()) label[pos[i]] = 1;
}

struct PredKeyEq {
    const uint32_t *key;
    uint32_t t;
    __device__ bool operator()(long long i) const { return key[i] == t; }
};

struct EmitFirstK {
    const int32_t *pos;
    int32_t *label;
    uint32_t k;
    __device__ void operator()(long long i, uint32_t r) const
    {
        if (r < k) label[pos[i]] = 1;
    }
};

template <typename IT>
uint32_t border_list(roift_ctx &c, const Dims &d, const void *img, const uint8_t *border, int32_t *pos,
                     uint32_t *key)
{
    const PredFlag pf{border};
    const uint32_t nb = cp_count(c, d.n, pf);
    cp_emit(c, d.n, pf, EmitBorder<IT>{(const IT *)img, pos, key});
    return nb;
}

}  // namespace

int run_sobel(roift_ctx &c, const Dims &d, const void *d_image, int dtype, int32_t *d_grad, bool normalize)
{
    ROIFT_REQUIRE(dtype == ROIFT_I16 || dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16 or I32");
    ROIFT_REQUIRE(d_image && d_grad, ROIFT_EINVAL, "null pointer");
    const dim3 grid(grid_for((d.n + SV - 1) / SV, 256, 4096));
    int *part = buf<int>(c, "sobel.part", 2 * (size_t)grid.x);
    int *maxes = buf<int>(c, "sobel.max", 2);
    ev_record(c, 14);
    if (dtype == ROIFT_I16)
        hipLaunchKernelGGL(k_sobel<int16_t>, grid, dim3(256), 0, c.stream, (const int16_t *)d_image, d, d_grad, part);
    else
        hipLaunchKernelGGL(k_sobel<int32_t>, grid, dim3(256), 0, c.stream, (const int32_t *)d_image, d, d_grad, part);
    ROIFT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_max_fold, dim3(1), dim3(1024), 0, c.stream, (const int *)part, (int)grid.x, maxes);
    ROIFT_HIP_CHECK(hipGetLastError());
    if (normalize) {
        hipLaunchKernelGGL(k_grad_normalize, grid, dim3(256), 0, c.stream, d_grad, d.n, (const int *)maxes);
        ROIFT_HIP_CHECK(hipGetLastError());
    }
    ev_record(c, 15);
    c.ktimes[5] = ev_ms(c, 14, 15);
    return ROIFT_OK;
}

int run_bfs_edt(roift_ctx &c, const Dims &d, const int32_t *d_label, const int32_t *d_seeds, int n_seeds,
                int32_t *d_dist)
{
    ROIFT_REQUIRE(d_label && d_dist && (d_seeds || n_seeds == 0) && n_seeds >= 0, ROIFT_EINVAL, "bad arguments");
    fill<int32_t>(c, d_dist, d.n, EDT_INF);
    if (n_seeds > 0) {
        hipLaunchKernelGGL(k_edt_sources, dim3(grid_for(n_seeds)), dim3(256), 0, c.stream, d_seeds, n_seeds, d_label,
                           d.n, d_dist);
        ROIFT_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_edt_rows, dim3(grid_for((long long)d.ys * d.zs * 64)), dim3(256), 0, c.stream, d_dist, d);
    hipLaunchKernelGGL(k_edt_pass<1>, dim3(grid_for((long long)d.xs * d.zs)), dim3(256), 0, c.stream, d_dist, d, false);
    hipLaunchKernelGGL(k_edt_pass<2>, dim3(grid_for((long long)d.xs * d.ys)), dim3(256), 0, c.stream, d_dist, d, true);
    ROIFT_HIP_CHECK(hipGetLastError());
    return ROIFT_OK;
}

int run_boundary_seeds(roift_ctx &c, const Dims &d, int32_t *d_label, int stride, bool geodesic,
                       const int32_t *d_seeds, int n_seeds, int32_t *d_out, int capacity, int *h_added)
{
    ROIFT_REQUIRE(stride > 0, ROIFT_EINVAL, "stride must be > 0");
    ROIFT_REQUIRE(d_label && (d_out || capacity == 0) && capacity >= 0, ROIFT_EINVAL, "bad arguments");
    const VoxelXYZ v{d.xs, d.ys, d.zs};
    uint32_t count;
    ev_record(c, 14);
    if (geodesic) {
        int32_t *dist = buf<int32_t>(c, "edt.dist", d.n);
        const int rc = run_bfs_edt(c, d, d_label, d_seeds, n_seeds, dist);
        if (rc != ROIFT_OK) return rc;
        const PredGeodesic pr{v, d_label, dist, stride};
        count = cp_count(c, d.n, pr);
        ROIFT_REQUIRE(count <= (uint32_t)capacity, ROIFT_EINVAL, "boundary seeds exceed the output capacity");
        cp_emit(c, d.n, pr, EmitSeed{d_label, d_out});
    } else {
        const PredStride pr{v, d_label, stride};
        count = cp_count(c, d.n, pr);
        ROIFT_REQUIRE(count <= (uint32_t)capacity, ROIFT_EINVAL, "boundary seeds exceed the output capacity");
        cp_emit(c, d.n, pr, EmitSeed{d_label, d_out});
    }
    ev_record(c, 15);
    c.ktimes[5] = ev_ms(c, 14, 15);
    if (h_added) *h_added = (int)count;
    return ROIFT_OK;
}

int run_dilation_conditional(roift_ctx &c, const Dims &d, float radius, const void *d_image, int dtype,
                             int32_t *d_label, float percentile, int *h_set)
{
    ROIFT_REQUIRE(dtype == ROIFT_I16 || dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16 or I32");
    ROIFT_REQUIRE(d_image && d_label, ROIFT_EINVAL, "null pointer");
    // Spheric(radius) in the reference's order (gft_adjrel3.cpp:47-85); the centre first
    DilAdj A{};
    const int r0 = (int)radius, r2 = (int)(radius * radius);
    A.n = 1;
    for (int dz = -r0; dz <= r0; dz++)
        for (int dy = -r0; dy <= r0; dy++)
            for (int dx = -r0; dx <= r0; dx++) {
                if (dx * dx + dy * dy + dz * dz > r2 || (dx == 0 && dy == 0 && dz == 0)) continue;
                ROIFT_REQUIRE(A.n < DIL_MAXN, ROIFT_EUNSUPPORTED, "dilation radius too large");
                A.dx[A.n] = dx;
                A.dy[A.n] = dy;
                A.dz[A.n] = dz;
                A.n++;
            }
    uint8_t *border = buf<uint8_t>(c, "dil.border", d.n);
    int32_t *pos = buf<int32_t>(c, "dil.pos", d.n);
    uint32_t *key = buf<uint32_t>(c, "dil.key", d.n);
    ev_record(c, 14);
    if (A.n == 7)
        hipLaunchKernelGGL(k_dil_border6, dim3((d.xs + DB_X - 1) / DB_X, (d.ys + DB_Y - 1) / DB_Y, (d.zs + DB_Z - 1) / DB_Z),
                           dim3(DB_X, DB_Y), 0, c.stream, (const int32_t *)d_label, d, border);
    else
        hipLaunchKernelGGL(k_dil_border, dim3(grid_for(d.n, 256, 8192)), dim3(256), 0, c.stream,
                           (const int32_t *)d_label, d, A, border);
    ROIFT_HIP_CHECK(hipGetLastError());
    const uint32_t nb = dtype == ROIFT_I16 ? border_list<int16_t>(c, d, d_image, border, pos, key)
                                           : border_list<int32_t>(c, d, d_image, border, pos, key);
    int set = 0;
    if (nb > 0) {
        // kept: the first (int)index + 1 entries of the border stably sorted by value
        // (oiftrelax.cpp:139-148): index = (float)((percentile / 100.0) * (count - 1))
        const float index_percentile = (float)(((double)percentile / 100.0) * (double)(nb - 1));
        const long long keep = (long long)(int)index_percentile + 1;
        if (keep >= (long long)nb) {   // the whole border
            hipLaunchKernelGGL(k_dil_all, dim3(grid_for(nb)), dim3(256), 0, c.stream, (const int32_t *)pos,
                               (long long)nb, d_label);
            ROIFT_HIP_CHECK(hipGetLastError());
            set = (int)nb;
        } else if (keep > 0) {
            // radix select of the keep-th smallest key, 8 bits at a time from the top
            uint32_t *hist = buf<uint32_t>(c, "dil.hist", 256);
            uint32_t prefix = 0, mask = 0, rem = (uint32_t)keep, h[256];
            for (int shift = 24; shift >= 0; shift -= 8) {
                fill<uint32_t>(c, hist, 256, 0u);
                hipLaunchKernelGGL(k_radix_hist, dim3(grid_for(nb, 256, 1024)), dim3(256), 0, c.stream,
                                   (const uint32_t *)key, (long long)nb, prefix, mask, shift, hist);
                ROIFT_HIP_CHECK(hipGetLastError());
                ROIFT_HIP_CHECK(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, c.stream));
                ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
                uint32_t digit = 0, cum = 0;
                while (digit < 255 && cum + h[digit] < rem) cum += h[digit++];
                rem -= cum;
                prefix |= digit << shift;
                mask |= 0xFFu << shift;
            }
            // every key below the threshold, then the first `rem` equal keys in raster order
            hipLaunchKernelGGL(k_dil_below, dim3(grid_for(nb)), dim3(256), 0, c.stream, (const uint32_t *)key,
                               (const int32_t *)pos, (long long)nb, prefix, d_label);
            ROIFT_HIP_CHECK(hipGetLastError());
            const PredKeyEq pe{key, prefix};
            cp_count(c, nb, pe);
            cp_emit(c, nb, pe, EmitFirstK{pos, d_label, rem});
            set = (int)keep;
        }
    }
    ev_record(c, 15);
    c.ktimes[5] = ev_ms(c, 14, 15);
    if (h_set) *h_set = set;
    return ROIFT_OK;
}

}  // namespace roift
