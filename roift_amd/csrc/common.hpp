// common.hpp -- shared device/host definitions for the gfx950 OIFT + relaxation kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/roift.h"

namespace roift {

constexpr uint32_t INF32 = 0xFFFFFFFFu;
constexpr uint64_t INF64 = 0xFFFFFFFFFFFFFFFFull;
constexpr int NIL = -1;

// Neighbour displacements in gft::AdjRel3::Spheric order (gft_adjrel3.cpp:47-85):
// raster dz, dy, dx over the ball with the centre moved to index 0.
// Spheric(1.0): -z, -y, -x, +x, +y, +z.  Spheric(1.75): the 26 neighbours in
// raster order.  Index i and n-i are opposite; indices > (n-1)/2 are "forward".
template <int NADJ> struct Adjacency;

template <> struct Adjacency<7> {
    static constexpr int n = 7, half = 3;
    __host__ __device__ static constexpr int dx(int i) { return i == 3 ? -1 : (i == 4 ? 1 : 0); }
    __host__ __device__ static constexpr int dy(int i) { return i == 2 ? -1 : (i == 5 ? 1 : 0); }
    __host__ __device__ static constexpr int dz(int i) { return i == 1 ? -1 : (i == 6 ? 1 : 0); }
};

template <> struct Adjacency<27> {
    static constexpr int n = 27, half = 13;
    // raster index r = (dz+1)*9 + (dy+1)*3 + (dx+1); i in 1..13 -> r = i-1, i in 14..26 -> r = i
    __host__ __device__ static constexpr int raster(int i) { return i <= 13 ? i - 1 : i; }
    __host__ __device__ static constexpr int dx(int i) { return raster(i) % 3 - 1; }
    __host__ __device__ static constexpr int dy(int i) { return (raster(i) / 3) % 3 - 1; }
    __host__ __device__ static constexpr int dz(int i) { return raster(i) / 9 - 1; }
};

struct Dims {
    int xs, ys, zs;
    long long n;
    // voxel spacing (gft's scn->dx/dy/dz): only the relaxation's Dpq and the Gaussian blur's
    // taps read it (gft_ift.cpp:8723-8731, gft_filtering3.cpp:106-126)
    float sx = 1.0f, sy = 1.0f, sz = 1.0f;
};

// Division by an invariant d for 0 <= n < 2^31 with one mul-hi and a shift
// (Granlund & Montgomery, "Division by invariant integers using multiplication", 1994).
struct FastDiv {
    uint32_t d = 1, m = 1;
    int s = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t dv) : d(dv)
    {
        while ((1ull << s) < dv) s++;
        m = (uint32_t)((((1ull << 32) * ((1ull << s) - dv)) / dv) + 1);
    }
    __host__ __device__ inline uint32_t div(uint32_t n) const
    {
#ifdef __HIP_DEVICE_COMPILE__
        return (__umulhi(n, m) + n) >> s;
#else
        return (uint32_t)((((uint64_t)n * m) >> 32) + n) >> s;
#endif
    }
};

// Voxel coordinates of a linear index (volumes below 2^31 voxels).
struct Coords {
    FastDiv fx, fy;
    int xs, ys, zs;
    explicit Coords(const Dims &d) : fx((uint32_t)d.xs), fy((uint32_t)d.ys), xs(d.xs), ys(d.ys), zs(d.zs) {}
    __device__ inline void of(uint32_t p, int &x, int &y, int &z) const
    {
        const uint32_t t = fx.div(p);
        x = (int)(p - t * (uint32_t)xs);
        z = (int)fy.div(t);
        y = (int)(t - (uint32_t)z * (uint32_t)ys);
    }
};

__host__ __device__ inline bool in_volume(const Dims &d, int x, int y, int z)
{
    return (unsigned)x < (unsigned)d.xs && (unsigned)y < (unsigned)d.ys && (unsigned)z < (unsigned)d.zs;
}

// Oriented edge weights: slot s of voxel p stores the edge to q = p + d_(half+1+s):
//   wa[s][p] = w_obj(p->q), wb[s][p] = w_obj(q->p);  w_bg(p->q) = w_obj(q->p).
// cls = 1 selects w_obj, cls = 0 selects w_bg.
template <int NADJ, typename WT>
__device__ inline uint32_t arc_weight(const WT *__restrict__ wa, const WT *__restrict__ wb, long long n, int i,
                                      long long p, long long q, int cls)
{
    using A = Adjacency<NADJ>;
    if (i > A::half) {
        const long long s = i - A::half - 1;
        return cls ? (uint32_t)wa[s * n + p] : (uint32_t)wb[s * n + p];
    } else {
        const long long s = (A::n - i) - A::half - 1;
        return cls ? (uint32_t)wb[s * n + q] : (uint32_t)wa[s * n + q];
    }
}

// Grid-stride helpers.
__device__ inline long long gtid() { return (long long)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline long long gstride() { return (long long)gridDim.x * blockDim.x; }

// Wave-aggregated append to a global list: one atomicAdd per wavefront.
__device__ inline uint32_t wave_append_index(uint32_t *counter, bool want)
{
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return INF32;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (!want) return INF32;
    const unsigned long long below = mask & ((1ull << lane) - 1ull);
    return base + (uint32_t)__popcll(below);
}

// Wave-aggregated append of `cnt` items per lane: returns this lane's first index in the
// list (one atomicAdd per wavefront).  Every lane of the wavefront must call it.
__device__ inline uint32_t wave_reserve(uint32_t *counter, uint32_t cnt)
{
    const int lane = __lane_id();
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, o);
        if (lane >= o) incl += v;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63);
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(counter, total);
    base = (uint32_t)__shfl((int)base, 0);
    return base + incl - cnt;
}

// Workgroup reduction of one value per thread (1-D workgroups of at most 1024 threads, a
// multiple of 64); the result is valid in thread 0.  Reductions publish one atomic per
// workgroup: device-scope atomics on one address from many wavefronts serialise (several
// ns each on this chip), so per-wavefront atomics on a shared counter cost far more than
// the reduction itself.
template <typename T, typename Op>
__device__ inline T block_reduce(T v, Op op)
{
    __shared__ T part[16];
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
    const int w = (int)threadIdx.x / 64, nw = (int)blockDim.x / 64;
    __syncthreads();   // part may still be read by a previous call
    if (__lane_id() == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < nw; i++) v = op(v, part[i]);
    return v;
}

struct OpMax {
    template <typename T> __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};
struct OpMin {
    template <typename T> __device__ T operator()(T a, T b) const { return a < b ? a : b; }
};
struct OpAdd {
    template <typename T> __device__ T operator()(T a, T b) const { return a + b; }
};

// Fill without hipMemsetAsync (kernels are safe inside captured graphs).
template <typename T>
__global__ void k_fill(T *__restrict__ a, long long n, T v)
{
    for (long long i = gtid(); i < n; i += gstride()) a[i] = v;
}

struct Error {
    int code;
    std::string msg;
};

}  // namespace roift

#define ROIFT_HIP_CHECK(expr)                                                                   \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            throw ::roift::Error{ROIFT_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
    } while (0)

#define ROIFT_REQUIRE(cond, code, text)                      \
    do {                                                     \
        if (!(cond)) throw ::roift::Error{(code), (text)};   \
    } while (0)
