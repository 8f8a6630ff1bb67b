// oift.hip -- exact parallel OIFT under f_max for gfx950.
//
// Reproduces gft::ift::OIFT_Multi / OIFT_Multi_PerClass (gft_ift.cpp:911-1083)
// bit for bit, FIFO tie-breaking of the Dial bucket queue included
// (gft_pqueue32.cpp:300-327, 446-451), without a priority queue:
//
// * The reference pops voxels in (value, insertion time) order; the cost
//   C(q) = max(C(pred q), value q) is the running flood level at q's pop and
//   never decreases in pop order.  All voxels of one cost "level" are settled
//   together.  Within a level, bucket `level` is a FIFO of entries; each entry's
//   pop triggers a sub-flood over arcs cheaper than the level (values < level
//   are popped at once), and offers equal to the level join the bucket tail as
//   the next "generation".  A step = one (level, generation).
// * Every unsettled voxel holds one 64-bit word  value<<32 | key  updated by
//   atomicMin.  The key orders offers in time: keys are issued per step in
//   the sorted order of the entries that claimed the offering voxel, so the
//   minimum word is the reference's first-inserted minimum offer.
// * In a step the entries, ranked by key, flood their label class's arcs
//   (< level) with min-rank propagation.  Two labels that reach one voxel
//   compete; all ranks below r_safe = min over such voxels of the larger of
//   the two classes' ranks are free of competition and are settled at once.
//   The rest repeat the flood.  Same-class labels share arcs, so the
//   minimum rank reaching a voxel is its sequential owner.
//
// Bound: latency of the level/flood sweeps, then HBM.  Algorithmic bytes per
// voxel: weights read once (2E values), label in/out, cost out.
#include "ctx.hpp"

namespace roift {
namespace {

enum : uint8_t { FL_FROZEN = 1, FL_SETTLED = 2 };

enum Ctr : int {
    C_LEVEL = 0,
    C_F0 = 1,      // frontier list A count
    C_F1 = 2,      // frontier list B count
    C_E0 = 3,      // entry list A count
    C_E1 = 4,      // entry list B count
    C_ND = 5,      // distinct keys this step
    C_W0 = 6,      // worklist counters (triple-buffered)
    C_W1 = 7,
    C_W2 = 8,
    C_T = 9,       // touched
    C_NS = 10,     // newly settled this step
    C_RSAFE = 11,
    C_LEFT = 12,
    C_ERR = 13,
    C_NEXTKEY = 14,
    C_KBASE = 15,
    C_NRUNS = 16,
    C_COUNT = 32
};

constexpr int SORT_MAX = 16384;

__device__ inline uint32_t ld_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct State {
    uint64_t *word;
    uint8_t *flags;
    uint32_t *run0, *run1;
    uint32_t *tstamp, *qstamp, *estamp, *keystamp;
    int32_t *keylab;
    uint8_t *keycls;
    uint32_t *rank_of_key;
    int32_t *ranklab;
    uint32_t *F[2], *E[2], *W[2], *T, *NS, *D;
    uint32_t *ctr;
    int32_t *label, *cost;
    const int8_t *lcls;
    int lcls_n, lcls_default;
};

__device__ inline int label_class(const State &s, int lb)
{
    return (lb >= 0 && lb < s.lcls_n) ? s.lcls[lb] : s.lcls_default;
}

__global__ void k_init(State s, long long n, long long nkeys)
{
    for (long long p = gtid(); p < n; p += gstride()) {
        s.word[p] = INF64;
        s.run0[p] = INF32;
        s.run1[p] = INF32;
        s.flags[p] = (s.label[p] != NIL) ? FL_FROZEN : 0;
        s.tstamp[p] = 0;
        s.qstamp[p] = 0;
        s.estamp[p] = 0;
        if (s.cost) s.cost[p] = INT32_MAX;
    }
    for (long long k = gtid(); k < nkeys; k += gstride()) s.keystamp[k] = 0;
    if (gtid() < C_COUNT) s.ctr[gtid()] = 0;
}

// validate seeds, unfreeze them and flag the start of each same-label run (S order)
__global__ void k_seed_check(State s, const int32_t *__restrict__ S, int ns, long long n, uint32_t *__restrict__ runflag)
{
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int v = S[i];
        uint32_t f = 0;
        if (v < 0 || v >= n) {
            atomicOr(&s.ctr[C_ERR], 1u);
        } else {
            if (s.label[v] == NIL) atomicOr(&s.ctr[C_ERR], 2u);
            else if (label_class(s, s.label[v]) < 0) atomicOr(&s.ctr[C_ERR], 8u);
            if (atomicAdd(&s.estamp[v], 1u) != 0) atomicOr(&s.ctr[C_ERR], 4u);
            s.flags[v] = 0;
            if (i > 0) {
                const int u = S[i - 1];
                const int lu = (u >= 0 && u < n) ? s.label[u] : NIL;
                f = (lu != s.label[v]) ? 1u : 0u;
            }
        }
        runflag[i] = f;
    }
}

// Inclusive scan, phase 1: per-block scan of 1024 elements, block totals out.
__global__ __launch_bounds__(1024) void k_scan_block(uint32_t *__restrict__ a, int n, uint32_t *__restrict__ tot)
{
    __shared__ uint32_t sh[1024];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    sh[threadIdx.x] = i < n ? a[i] : 0;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        uint32_t v = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    if (i < n) a[i] = sh[threadIdx.x];
    if (threadIdx.x == 1023) tot[blockIdx.x] = sh[1023];
}

// phase 2: one block turns the block totals into exclusive offsets
__global__ __launch_bounds__(1024) void k_scan_totals(uint32_t *__restrict__ tot, int nb)
{
    __shared__ uint32_t sh[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const uint32_t x = i < nb ? tot[i] : 0;
        sh[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            uint32_t v = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < nb) tot[i] = carry + sh[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += sh[1023];
        __syncthreads();
    }
}

// seeds: key = run index (consecutive same-label seeds share a key), value 0
__global__ void k_seed_init(State s, const int32_t *__restrict__ S, int ns, long long n,
                            const uint32_t *__restrict__ runinc, const uint32_t *__restrict__ tot)
{
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int v = S[i];
        if (v < 0 || v >= n) continue;  // reported by k_seed_check; the host stops before any use
        const uint32_t key = runinc[i] + tot[i / 1024];
        s.word[v] = (uint64_t)key;
        const int lb = s.label[v];
        s.keylab[key] = lb;
        s.keycls[key] = (uint8_t)label_class(s, lb);
        s.F[0][i] = (uint32_t)v;
        if (i == ns - 1) {
            s.ctr[C_NEXTKEY] = key + 1;
            s.ctr[C_NRUNS] = key + 1;
            s.ctr[C_F0] = (uint32_t)ns;
        }
    }
}

__global__ void k_set(uint32_t *ctr, int idx, uint32_t v)
{
    if (gtid() == 0) ctr[idx] = v;
}

__global__ void k_set2(uint32_t *ctr, int i0, uint32_t v0, int i1, uint32_t v1)
{
    if (gtid() == 0) {
        ctr[i0] = v0;
        ctr[i1] = v1;
    }
}

// Compact the frontier (drop settled voxels) and find the minimum pending value.
__global__ __launch_bounds__(256) void k_level_scan(State s, int fin)
{
    const uint32_t *in = s.F[fin];
    uint32_t *out = s.F[fin ^ 1];
    const uint32_t nin = s.ctr[C_F0 + fin];
    uint32_t lmin = INF32;
    for (long long base = (long long)blockIdx.x * blockDim.x; base < nin; base += gstride()) {
        const long long i = base + threadIdx.x;
        bool keep = false;
        uint32_t v = 0;
        if (i < nin) {
            v = in[i];
            if (!(s.flags[v] & FL_SETTLED)) {
                keep = true;
                lmin = min(lmin, (uint32_t)(s.word[v] >> 32));
            }
        }
        const uint32_t idx = wave_append_index(&s.ctr[C_F0 + (fin ^ 1)], keep);
        if (keep) out[idx] = v;
    }
    for (int o = 32; o > 0; o >>= 1) lmin = min(lmin, (uint32_t)__shfl_xor((int)lmin, o));
    if (__lane_id() == 0 && lmin != INF32) atomicMin(&s.ctr[C_LEVEL], lmin);
}

__global__ __launch_bounds__(256) void k_collect(State s, int fin, int eout)
{
    const uint32_t *in = s.F[fin];
    const uint32_t nin = s.ctr[C_F0 + fin];
    const uint32_t level = s.ctr[C_LEVEL];
    for (long long base = (long long)blockIdx.x * blockDim.x; base < nin; base += gstride()) {
        const long long i = base + threadIdx.x;
        bool take = false;
        uint32_t v = 0;
        if (i < nin) {
            v = in[i];
            take = (uint32_t)(s.word[v] >> 32) == level;
        }
        const uint32_t idx = wave_append_index(&s.ctr[C_E0 + eout], take);
        if (take) s.E[eout][idx] = v;
    }
}

// distinct keys of the step's entries
__global__ __launch_bounds__(256) void k_rank_collect(State s, int ein, uint32_t step)
{
    const uint32_t *E = s.E[ein];
    const uint32_t ne = s.ctr[C_E0 + ein];
    for (long long base = (long long)blockIdx.x * blockDim.x; base < ne; base += gstride()) {
        const long long i = base + threadIdx.x;
        bool fresh = false;
        uint32_t key = 0;
        if (i < ne) {
            key = (uint32_t)s.word[E[i]];
            fresh = atomicExch(&s.keystamp[key], step) != step;
        }
        const uint32_t idx = wave_append_index(&s.ctr[C_ND], fresh);
        if (fresh) s.D[idx] = key;
    }
}

__device__ inline void publish_rank(State &s, uint32_t key, uint32_t r, uint32_t base)
{
    s.rank_of_key[key] = r;
    const int lb = s.keylab[key];
    s.ranklab[r] = lb;
    s.keylab[base + r] = lb;
    s.keycls[base + r] = s.keycls[key];
}

// Sort the distinct keys (bitonic in LDS) and issue this step's key block.
__global__ __launch_bounds__(1024) void k_rank_sort(State s)
{
    __shared__ uint32_t sh[SORT_MAX];
    const uint32_t nd = s.ctr[C_ND];
    if (nd > (uint32_t)SORT_MAX) return;  // k_rank_count handles large key sets
    const uint32_t base = s.ctr[C_NEXTKEY];
    uint32_t P = 1;
    while (P < nd) P <<= 1;
    for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) sh[t] = t < nd ? s.D[t] : INF32;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
                const uint32_t u = t ^ j;
                if (u > t) {
                    const uint32_t a = sh[t], b = sh[u];
                    const bool up = (t & k) == 0;
                    if ((a > b) == up) {
                        sh[t] = b;
                        sh[u] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t t = threadIdx.x; t < nd; t += blockDim.x) publish_rank(s, sh[t], t, base);
}

// Fallback for large key sets: rank = number of distinct keys below the key.
__global__ __launch_bounds__(256) void k_rank_count(State s)
{
    __shared__ uint32_t tile[2048];
    const uint32_t nd = s.ctr[C_ND];
    if (nd <= (uint32_t)SORT_MAX) return;  // sorted by k_rank_sort
    const uint32_t base = s.ctr[C_NEXTKEY];
    for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < nd; b0 += gstride()) {
        const long long i = b0 + threadIdx.x;
        const uint32_t key = i < nd ? s.D[i] : INF32;
        uint32_t r = 0;
        for (uint32_t t0 = 0; t0 < nd; t0 += 2048) {
            __syncthreads();
            for (uint32_t t = threadIdx.x; t < 2048; t += blockDim.x) tile[t] = (t0 + t < nd) ? s.D[t0 + t] : INF32;
            __syncthreads();
            for (int t = 0; t < 2048; t++) r += tile[t] < key ? 1u : 0u;
        }
        if (i < nd) publish_rank(s, key, r, base);
    }
}

__global__ void k_rank_finish(State s)
{
    if (gtid() == 0) {
        const uint32_t nd = s.ctr[C_ND];
        const uint32_t base = s.ctr[C_NEXTKEY];
        s.ctr[C_KBASE] = base;
        s.ctr[C_NEXTKEY] = base + nd;
    }
}

// Start a flood iteration from the entries that are still unsettled.
__global__ __launch_bounds__(256) void k_iter_seed(State s, int ein, uint32_t iter, uint32_t round, int wout)
{
    const uint32_t *E = s.E[ein];
    const uint32_t ne = s.ctr[C_E0 + ein];
    for (long long base = (long long)blockIdx.x * blockDim.x; base < ne; base += gstride()) {
        const long long i = base + threadIdx.x;
        bool add = false;
        uint32_t v = 0;
        if (i < ne) {
            v = E[i];
            if (!(s.flags[v] & FL_SETTLED)) {
                const uint32_t key = (uint32_t)s.word[v];
                const uint32_t r = s.rank_of_key[key];
                if (s.keycls[key]) s.run1[v] = r;
                else s.run0[v] = r;
                s.tstamp[v] = iter;
                s.qstamp[v] = round;
                add = true;
            }
        }
        const uint32_t it = wave_append_index(&s.ctr[C_T], add);
        if (add) s.T[it] = v;
        const uint32_t iw = wave_append_index(&s.ctr[C_W0 + wout], add);
        if (add) s.W[0][iw] = v;
    }
}

// One flood round: min-rank propagation along arcs cheaper than the level.
template <int NADJ, typename WT>
__global__ __launch_bounds__(256) void k_flood(State s, const WT *__restrict__ wa, const WT *__restrict__ wb, Dims d,
                                                int lin, int cin, int cout, int czero, uint32_t iter, uint32_t round)
{
    using A = Adjacency<NADJ>;
    if (gtid() == 0) s.ctr[C_W0 + czero] = 0;
    const uint32_t *in = s.W[lin];
    uint32_t *out = s.W[lin ^ 1];
    const uint32_t nin = s.ctr[C_W0 + cin];
    const uint32_t level = s.ctr[C_LEVEL];
    const long long xy = (long long)d.xs * d.ys;
    for (long long base = (long long)blockIdx.x * blockDim.x; base < nin; base += gstride()) {
        const long long i = base + threadIdx.x;
        const bool valid = i < nin;
        long long p = 0;
        uint32_t r0 = INF32, r1 = INF32;
        int x = 0, y = 0, z = 0;
        if (valid) {
            p = in[i];
            r0 = ld_relaxed(&s.run0[p]);
            r1 = ld_relaxed(&s.run1[p]);
            x = (int)(p % d.xs);
            y = (int)((p / d.xs) % d.ys);
            z = (int)(p / xy);
        }
#pragma unroll
        for (int k = 1; k < NADJ; k++) {
            bool push = false;
            long long q = 0;
            const int qx = x + A::dx(k), qy = y + A::dy(k), qz = z + A::dz(k);
            if (valid && in_volume(d, qx, qy, qz)) {
                q = qx + (long long)qy * d.xs + (long long)qz * xy;
                if (!(s.flags[q] & (FL_FROZEN | FL_SETTLED))) {
                    if (r0 != INF32 && arc_weight<NADJ, WT>(wa, wb, d.n, k, p, q, 0) < level) {
                        if (atomicMin(&s.run0[q], r0) > r0) push = true;
                    }
                    if (r1 != INF32 && arc_weight<NADJ, WT>(wa, wb, d.n, k, p, q, 1) < level) {
                        if (atomicMin(&s.run1[q], r1) > r1) push = true;
                    }
                }
            }
            bool tadd = false, qadd = false;
            if (push) {
                tadd = atomicExch(&s.tstamp[q], iter) != iter;
                qadd = atomicExch(&s.qstamp[q], round) != round;
            }
            const uint32_t it = wave_append_index(&s.ctr[C_T], tadd);
            if (tadd) s.T[it] = (uint32_t)q;
            const uint32_t iq = wave_append_index(&s.ctr[C_W0 + cout], qadd);
            if (qadd) out[iq] = (uint32_t)q;
        }
    }
}

__global__ __launch_bounds__(256) void k_rsafe(State s)
{
    const uint32_t nt = s.ctr[C_T];
    uint32_t lmin = INF32;
    for (long long i = gtid(); i < nt; i += gstride()) {
        const uint32_t v = s.T[i];
        const uint32_t a = s.run0[v], b = s.run1[v];
        if (a != INF32 && b != INF32) lmin = min(lmin, max(a, b));
    }
    for (int o = 32; o > 0; o >>= 1) lmin = min(lmin, (uint32_t)__shfl_xor((int)lmin, o));
    if (__lane_id() == 0 && lmin != INF32) atomicMin(&s.ctr[C_RSAFE], lmin);
}

__global__ __launch_bounds__(256) void k_claim(State s)
{
    const uint32_t nt = s.ctr[C_T];
    const uint32_t rs = s.ctr[C_RSAFE];
    const uint32_t level = s.ctr[C_LEVEL];
    for (long long base = (long long)blockIdx.x * blockDim.x; base < nt; base += gstride()) {
        const long long i = base + threadIdx.x;
        bool settle = false;
        uint32_t v = 0;
        if (i < nt) {
            v = s.T[i];
            const uint32_t a = s.run0[v], b = s.run1[v];
            if (a < rs) {
                settle = true;
                s.run1[v] = INF32;
                s.label[v] = s.ranklab[a];
            } else if (b < rs) {
                settle = true;
                s.run0[v] = INF32;
                s.label[v] = s.ranklab[b];
            } else {
                s.run0[v] = INF32;
                s.run1[v] = INF32;
            }
            if (settle) {
                s.flags[v] = FL_SETTLED;
                if (s.cost) s.cost[v] = (int32_t)level;
            }
        }
        const uint32_t idx = wave_append_index(&s.ctr[C_NS], settle);
        if (settle) s.NS[idx] = v;
    }
}

__global__ __launch_bounds__(256) void k_count_left(State s, int ein)
{
    const uint32_t *E = s.E[ein];
    const uint32_t ne = s.ctr[C_E0 + ein];
    uint32_t c = 0;
    for (long long i = gtid(); i < ne; i += gstride()) c += (s.flags[E[i]] & FL_SETTLED) ? 0u : 1u;
    for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o);
    if (__lane_id() == 0 && c) atomicAdd(&s.ctr[C_LEFT], c);
}

// Offers from the voxels settled in this step: word = min(word, w<<32 | key).
template <int NADJ, typename WT>
__global__ __launch_bounds__(256) void k_offers(State s, const WT *__restrict__ wa, const WT *__restrict__ wb, Dims d,
                                                 int fout, int eout)
{
    using A = Adjacency<NADJ>;
    const uint32_t nns = s.ctr[C_NS];
    const uint32_t level = s.ctr[C_LEVEL];
    const uint32_t kbase = s.ctr[C_KBASE];
    const long long xy = (long long)d.xs * d.ys;
    for (long long base = (long long)blockIdx.x * blockDim.x; base < nns; base += gstride()) {
        const long long i = base + threadIdx.x;
        const bool valid = i < nns;
        long long p = 0;
        int cls = 0, x = 0, y = 0, z = 0;
        uint64_t K = 0;
        if (valid) {
            p = s.NS[i];
            const uint32_t a = s.run0[p], b = s.run1[p];
            cls = (b != INF32) ? 1 : 0;
            K = (uint64_t)(kbase + (cls ? b : a));
            x = (int)(p % d.xs);
            y = (int)((p / d.xs) % d.ys);
            z = (int)(p / xy);
        }
#pragma unroll
        for (int k = 1; k < NADJ; k++) {
            bool fadd = false, eadd = false;
            long long q = 0;
            const int qx = x + A::dx(k), qy = y + A::dy(k), qz = z + A::dz(k);
            if (valid && in_volume(d, qx, qy, qz)) {
                q = qx + (long long)qy * d.xs + (long long)qz * xy;
                if (!(s.flags[q] & (FL_FROZEN | FL_SETTLED))) {
                    const uint32_t w = arc_weight<NADJ, WT>(wa, wb, d.n, k, p, q, cls);
                    const uint64_t nw = ((uint64_t)w << 32) | K;
                    const uint64_t old = atomicMin((unsigned long long *)&s.word[q], (unsigned long long)nw);
                    if (nw < old && (uint32_t)(old >> 32) > w) {
                        fadd = (old == INF64);
                        eadd = (w == level);
                    }
                }
            }
            const uint32_t jf = wave_append_index(&s.ctr[C_F0 + fout], fadd);
            if (fadd) s.F[fout][jf] = (uint32_t)q;
            const uint32_t je = wave_append_index(&s.ctr[C_E0 + eout], eadd);
            if (eadd) s.E[eout][je] = (uint32_t)q;
        }
    }
}

__global__ void k_reset_runs(State s)
{
    const uint32_t nns = s.ctr[C_NS];
    for (long long i = gtid(); i < nns; i += gstride()) {
        const uint32_t v = s.NS[i];
        s.run0[v] = INF32;
        s.run1[v] = INF32;
    }
}

template <int NADJ, typename WT>
int drive(roift_ctx &c, State &s, const Dims &d, const WT *wa, const WT *wb, const int32_t *d_seeds, int ns)
{
    hipStream_t st = c.stream;
    uint32_t *h = c.h_ctr;
    const long long n = d.n;
    const long long nkeys = 2 * n + 1024;
    auto sync_read = [&](int first, int count) {
        ROIFT_HIP_CHECK(hipMemcpyAsync(h + first, s.ctr + first, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, st));
        ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    };
    const unsigned G = grid_for(n, 256, 8192);
    hipLaunchKernelGGL(k_init, dim3(G), dim3(256), 0, st, s, n, nkeys);
    ROIFT_HIP_CHECK(hipGetLastError());
    if (ns == 0) return 0;
    uint32_t *runflag = buf<uint32_t>(c, "oift.runflag", (size_t)ns);
    const int nb = (ns + 1023) / 1024;
    uint32_t *tot = buf<uint32_t>(c, "oift.runtot", (size_t)nb);
    hipLaunchKernelGGL(k_seed_check, dim3(grid_for(ns)), dim3(256), 0, st, s, d_seeds, ns, n, runflag);
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(1024), 0, st, runflag, ns, tot);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(1024), 0, st, tot, nb);
    hipLaunchKernelGGL(k_seed_init, dim3(grid_for(ns)), dim3(256), 0, st, s, d_seeds, ns, n, runflag, tot);
    ROIFT_HIP_CHECK(hipGetLastError());
    sync_read(C_ERR, 1);
    if (h[C_ERR]) {
        const uint32_t e = h[C_ERR];
        if (e & 8) throw Error{ROIFT_EUNSUPPORTED, "seed label outside per_class with polarity 0 next to +-P"};
        throw Error{ROIFT_EINVAL, (e & 1) ? "seed index out of range"
                                          : (e & 2) ? "seed with NIL label" : "duplicate seed index"};
    }

    int fin = 0;          // live frontier list
    uint32_t step = 0, iter = 0, round = 0;
    long long nlevels = 0, nrounds = 0;
    for (;;) {
        // next level: min pending value; compact the frontier into the other list
        hipLaunchKernelGGL(k_set2, dim3(1), dim3(64), 0, st, s.ctr, C_LEVEL, INF32, C_F0 + (fin ^ 1), 0u);
        hipLaunchKernelGGL(k_level_scan, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, s, fin);
        fin ^= 1;
        hipLaunchKernelGGL(k_set, dim3(1), dim3(64), 0, st, s.ctr, C_E0, 0u);
        hipLaunchKernelGGL(k_collect, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, s, fin, 0);
        ROIFT_HIP_CHECK(hipGetLastError());
        sync_read(0, C_COUNT);
        if (h[C_LEVEL] == INF32) break;
        nlevels++;
        int ein = 0;
        for (;;) {  // generations of this level
            step++;
            const uint32_t ne = h[C_E0 + ein];
            hipLaunchKernelGGL(k_set2, dim3(1), dim3(64), 0, st, s.ctr, C_ND, 0u, C_NS, 0u);
            hipLaunchKernelGGL(k_rank_collect, dim3(grid_for(ne, 256, 4096)), dim3(256), 0, st, s, ein, step);
            hipLaunchKernelGGL(k_rank_sort, dim3(1), dim3(1024), 0, st, s);
            hipLaunchKernelGGL(k_rank_count, dim3(256), dim3(256), 0, st, s);
            hipLaunchKernelGGL(k_rank_finish, dim3(1), dim3(64), 0, st, s);
            ROIFT_HIP_CHECK(hipGetLastError());
            for (;;) {  // r_safe iterations
                iter++;
                round++;
                hipLaunchKernelGGL(k_set2, dim3(1), dim3(64), 0, st, s.ctr, C_T, 0u, C_W0, 0u);
                hipLaunchKernelGGL(k_set, dim3(1), dim3(64), 0, st, s.ctr, C_W1, 0u);
                hipLaunchKernelGGL(k_set2, dim3(1), dim3(64), 0, st, s.ctr, C_RSAFE, INF32, C_LEFT, 0u);
                hipLaunchKernelGGL(k_iter_seed, dim3(grid_for(ne, 256, 4096)), dim3(256), 0, st, s, ein, iter,
                                   round, 0);
                ROIFT_HIP_CHECK(hipGetLastError());
                // flood rounds: worklist counters rotate W0 -> W1 -> W2, lists alternate
                int r = 0;
                uint32_t approx = ne;
                for (;;) {
                    const int batch = 4;
                    for (int b = 0; b < batch; b++, r++) {
                        round++;
                        const int cin = r % 3, cout = (r + 1) % 3, czero = (r + 2) % 3;
                        hipLaunchKernelGGL((k_flood<NADJ, WT>), dim3(grid_for(approx, 256, 4096)), dim3(256), 0, st,
                                           s, wa, wb, d, r & 1, cin, cout, czero, iter, round);
                    }
                    ROIFT_HIP_CHECK(hipGetLastError());
                    nrounds += batch;
                    sync_read(C_W0, 3);
                    approx = h[C_W0 + r % 3];
                    if (approx == 0) break;
                }
                hipLaunchKernelGGL(k_rsafe, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, s);
                hipLaunchKernelGGL(k_claim, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, s);
                hipLaunchKernelGGL(k_count_left, dim3(grid_for(ne, 256, 1024)), dim3(256), 0, st, s, ein);
                ROIFT_HIP_CHECK(hipGetLastError());
                sync_read(C_LEFT, 1);
                if (h[C_LEFT] == 0) break;
            }
            // offers and the next generation
            hipLaunchKernelGGL(k_set, dim3(1), dim3(64), 0, st, s.ctr, C_E0 + (ein ^ 1), 0u);
            hipLaunchKernelGGL((k_offers<NADJ, WT>), dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, s, wa, wb, d,
                               fin, ein ^ 1);
            hipLaunchKernelGGL(k_reset_runs, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, s);
            ROIFT_HIP_CHECK(hipGetLastError());
            sync_read(0, C_COUNT);
            ein ^= 1;
            if (h[C_E0 + ein] == 0) break;
        }
    }
    c.timings[4] = (double)step;
    c.timings[5] = (double)nrounds;
    return (int)nlevels;
}

}  // namespace

int run_oift(roift_ctx &c, const Dims &d, int nadj, const void *d_wa, const void *d_wb, int weight_dtype,
             const LabelClasses &lc, const int32_t *d_seeds, int n_seeds, int32_t *d_label, int32_t *d_cost)
{
    ROIFT_REQUIRE(nadj == 7 || nadj == 27, ROIFT_EUNSUPPORTED, "adjacency must be Spheric(1.0) or Spheric(1.75)");
    ROIFT_REQUIRE(d.n > 0 && d.n < (1ll << 31), ROIFT_EINVAL, "volume must hold 1 .. 2^31-1 voxels");
    ROIFT_REQUIRE(n_seeds >= 0 && n_seeds <= d.n, ROIFT_EINVAL, "bad seed count");
    const long long n = d.n;
    State s{};
    s.word = buf<uint64_t>(c, "oift.word", n);
    s.flags = buf<uint8_t>(c, "oift.flags", n);
    s.run0 = buf<uint32_t>(c, "oift.run0", n);
    s.run1 = buf<uint32_t>(c, "oift.run1", n);
    s.tstamp = buf<uint32_t>(c, "oift.tstamp", n);
    s.qstamp = buf<uint32_t>(c, "oift.qstamp", n);
    s.estamp = buf<uint32_t>(c, "oift.estamp", n);
    const long long nkeys = 2 * n + 1024;
    s.keystamp = buf<uint32_t>(c, "oift.keystamp", nkeys);
    s.keylab = buf<int32_t>(c, "oift.keylab", nkeys);
    s.keycls = buf<uint8_t>(c, "oift.keycls", nkeys);
    s.rank_of_key = buf<uint32_t>(c, "oift.rank_of_key", nkeys);
    s.ranklab = buf<int32_t>(c, "oift.ranklab", n + 1024);
    s.F[0] = buf<uint32_t>(c, "oift.F0", n);
    s.F[1] = buf<uint32_t>(c, "oift.F1", n);
    s.E[0] = buf<uint32_t>(c, "oift.E0", n);
    s.E[1] = buf<uint32_t>(c, "oift.E1", n);
    s.W[0] = buf<uint32_t>(c, "oift.W0", n);
    s.W[1] = buf<uint32_t>(c, "oift.W1", n);
    s.T = buf<uint32_t>(c, "oift.T", n);
    s.NS = buf<uint32_t>(c, "oift.NS", n);
    s.D = buf<uint32_t>(c, "oift.D", n);
    s.ctr = buf<uint32_t>(c, "oift.ctr", C_COUNT);
    s.label = d_label;
    s.cost = d_cost;
    const int ncls = (int)lc.cls.size();
    int8_t *lcls = buf<int8_t>(c, "oift.lcls", ncls > 0 ? ncls : 1);
    if (ncls > 0)
        ROIFT_HIP_CHECK(hipMemcpyAsync(lcls, lc.cls.data(), ncls, hipMemcpyHostToDevice, c.stream));
    s.lcls = lcls;
    s.lcls_n = ncls;
    s.lcls_default = lc.default_cls;
    if (weight_dtype == ROIFT_U16) {
        if (nadj == 7) return drive<7, uint16_t>(c, s, d, (const uint16_t *)d_wa, (const uint16_t *)d_wb, d_seeds, n_seeds);
        return drive<27, uint16_t>(c, s, d, (const uint16_t *)d_wa, (const uint16_t *)d_wb, d_seeds, n_seeds);
    }
    ROIFT_REQUIRE(weight_dtype == ROIFT_U32, ROIFT_EINVAL, "weight dtype must be U16 or U32");
    if (nadj == 7) return drive<7, uint32_t>(c, s, d, (const uint32_t *)d_wa, (const uint32_t *)d_wb, d_seeds, n_seeds);
    return drive<27, uint32_t>(c, s, d, (const uint32_t *)d_wa, (const uint32_t *)d_wb, d_seeds, n_seeds);
}

}  // namespace roift
