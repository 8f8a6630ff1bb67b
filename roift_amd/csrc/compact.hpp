// compact.hpp -- order-preserving selection over an index range [0, n).
//
// Two passes over the same predicate: k_cp_count counts the selected indices of each
// 4096-index tile, k_cp_scan turns the counts into tile offsets, and k_cp_emit calls
// emit(p, rank) for every selected p, where rank is p's position among the selected
// indices in increasing order.  This is how the seed lists and the dilation border keep
// the reference's raster order (its loops append in that order, and seed order is the
// OIFT's FIFO tie order).
#pragma once

#include "ctx.hpp"

namespace roift {

constexpr int CP_THREADS = 256, CP_ITEMS = 16, CP_TILE = CP_THREADS * CP_ITEMS;

// Exclusive scan of one value per thread over a workgroup of NT threads; `total` receives
// the workgroup's sum.  Every thread must call it (it synchronises).
template <int NT>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t &total)
{
    __shared__ uint32_t wsum[NT / 64];
    const int lane = __lane_id(), w = (int)threadIdx.x / 64;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) {
        if (i < w) base += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();   // wsum may be reused by the next call
    total = tot;
    return base + incl - v;
}

template <class Pred>
__global__ __launch_bounds__(CP_THREADS) void k_cp_count(Pred pr, long long n, uint32_t *__restrict__ blk)
{
    const long long base = (long long)blockIdx.x * CP_TILE + (long long)threadIdx.x * CP_ITEMS;
    uint32_t c = 0;
    for (int j = 0; j < CP_ITEMS; j++) {
        const long long p = base + j;
        if (p < n && pr(p)) c++;
    }
    uint32_t total;
    block_exclusive_scan<CP_THREADS>(c, total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// In-place exclusive scan of nb tile counts by one workgroup of 1024 threads; the grand
// total goes to blk[nb].
__global__ __launch_bounds__(1024) void k_cp_scan(uint32_t *__restrict__ blk, int nb);

template <class Pred, class Emit>
__global__ __launch_bounds__(CP_THREADS) void k_cp_emit(Pred pr, Emit em, long long n,
                                                        const uint32_t *__restrict__ blk)
{
    const long long base = (long long)blockIdx.x * CP_TILE + (long long)threadIdx.x * CP_ITEMS;
    uint32_t flags = 0, c = 0;
    for (int j = 0; j < CP_ITEMS; j++) {
        const long long p = base + j;
        if (p < n && pr(p)) {
            flags |= 1u << j;
            c++;
        }
    }
    uint32_t total;
    uint32_t r = blk[blockIdx.x] + block_exclusive_scan<CP_THREADS>(c, total);
    for (int j = 0; j < CP_ITEMS; j++)
        if ((flags >> j) & 1u) em(base + j, r++);
}

// Host side: count, then (optionally) emit with the same predicate.
inline uint32_t *cp_blocks(roift_ctx &c, long long n, int &nb)
{
    nb = (int)((n + CP_TILE - 1) / CP_TILE);
    if (nb < 1) nb = 1;
    return buf<uint32_t>(c, "cp.blk", (size_t)nb + 1);
}

template <class Pred>
uint32_t cp_count(roift_ctx &c, long long n, const Pred &pr)
{
    int nb;
    uint32_t *blk = cp_blocks(c, n, nb);
    hipLaunchKernelGGL((k_cp_count<Pred>), dim3(nb), dim3(CP_THREADS), 0, c.stream, pr, n, blk);
    ROIFT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_cp_scan, dim3(1), dim3(1024), 0, c.stream, blk, nb);
    ROIFT_HIP_CHECK(hipGetLastError());
    uint32_t total = 0;
    ROIFT_HIP_CHECK(hipMemcpyAsync(&total, blk + nb, sizeof(total), hipMemcpyDeviceToHost, c.stream));
    ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
    return total;
}

// Must follow cp_count with the same n and an equivalent predicate.
template <class Pred, class Emit>
void cp_emit(roift_ctx &c, long long n, const Pred &pr, const Emit &em)
{
    int nb;
    uint32_t *blk = cp_blocks(c, n, nb);
    hipLaunchKernelGGL((k_cp_emit<Pred, Emit>), dim3(nb), dim3(CP_THREADS), 0, c.stream, pr, em, n, blk);
    ROIFT_HIP_CHECK(hipGetLastError());
}

}  // namespace roift
