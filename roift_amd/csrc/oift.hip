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
// * In a step the entries, ranked by key, flood their polarity class's arcs
//   (< level) with min-rank propagation.  Two classes that reach one voxel
//   compete; every rank below r_safe = min over such voxels of the larger of
//   the two classes' ranks is free of competition and is settled at once.
//   The rest repeat the flood.  Same-class labels share arcs, so the minimum
//   rank reaching a voxel is its sequential owner.
//
// The whole level/step/iteration/round loop runs in ONE persistent cooperative
// kernel: phases are separated by grid barriers and every control decision is
// taken by all workgroups from counters read right after a barrier.  Counters
// that steer the loop live in parity/round-indexed slots and are only reset
// after a further barrier, so no workgroup can see a reset before deciding.
// Inside a flood round each workgroup keeps the voxels it discovers in an LDS
// queue and relaxes them at once (min-propagation converges in any order);
// only LDS overflow and work beyond a per-round budget go to the next round.
//
// Bound: latency of the barrier-separated phases; algorithmic bytes per voxel:
// weights read once (2E values), label in/out, cost out.
#include <cmath>

#include "ctx.hpp"

namespace roift {
namespace {

enum : uint8_t { FL_FROZEN = 1, FL_SETTLED = 2 };

// Counter slots, laid out by 128-byte cache line (32 words).  The words that waiting workgroups
// poll (the barrier generation and the abort flag), the flood's work cursors, the barrier's
// top counter and each group counter sit on lines of their own: polls and atomics on one line
// are served one after another, and the workgroups that still work in a phase would otherwise
// queue behind the ones spinning at its end.  The first K_HOST words are copied to the host.
enum Ctr : int {
    K_LEVEL = 0,    // [2] by level parity
    K_F = 2,        // [2] frontier list counts
    K_E = 4,        // [2] entry list counts
    K_ND = 6,       // distinct keys of the step
    K_NS = 7,       // voxels settled in the step
    K_T = 8,        // [2] touched, by iteration parity
    K_WC = 10,      // [3] worklist counts, by round mod 3
    K_KW0 = 13,     // [2] first bitmap word of the step's keys, by step parity
    K_RSAFE = 16,   // [2] by iteration parity
    K_LEFT = 18,    // [2] by iteration parity
    K_ERR = 20,
    K_NEXTKEY = 21,
    K_KBASE = 22,
    K_NLEVELS = 23,
    K_NSTEPS = 24,
    K_NITERS = 25,
    K_NROUNDS = 26,
    K_FIN = 27,     // which frontier list is live when the kernel starts
    K_CONT = 28,    // [2] "two classes met" flag, by iteration parity
    K_KW1 = 30,     // [2] last bitmap word
    K_ABORT = 32,   // line 1: polled by waiting workgroups
    K_BAR_GEN = 33,
    K_HOST = 64,    // words mirrored on the host
    K_CUR = 64,     // line 2: [3] worklist cursors, by round mod 3
    K_RMAXB = 68,   // [2] u64 by round parity: the round's busiest workgroup, batches << 32 | items (profile)
    K_CLAIMED = 72, // [2] by iteration parity: the step was claimed inside the flood's last barrier
    K_BAR_ARRIVE = 96,    // line 3
    K_BAR_GROUP = 128,    // lines 4..: [BAR_GROUPS] first-level barrier counters, K_LINE apart
    K_LINE = 32,
    K_COUNT = 128 + 16 * 32
};

constexpr int BLOCK = 512;
constexpr int SORT_MAX = 16384;     // LDS words: bitonic sort capacity and flood queue
constexpr int QCAP = SORT_MAX;
constexpr int QUAD_MAX = 64;        // steps with at most this many entries rank keys in one wavefront
constexpr uint32_t BAR_GROUPS = 16;
static_assert(K_BAR_GROUP + (int)BAR_GROUPS * K_LINE <= K_COUNT, "barrier group lines fit the counters");
constexpr int MAX_BLOCKS = 1024;   // persistent grid cap (bounds the touched list's chunk slack)

__device__ inline uint32_t ld(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void st(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct State {
    // 16-bit weights: every voxel's arc weights of both classes in one slot (u32 k-1 =
    // class-1 weight << 16 | class-0 weight of arc k; 32 bytes for 6-adjacency, 128 for
    // 26-adjacency); null for 32-bit weights
    const uint32_t *wpk;
    uint64_t *word;
    uint8_t *flags;
    uint64_t *run;    // per-class minimum rank of the current flood: class 0 low, class 1 high
    uint32_t *qstamp, *estamp, *keystamp;
    int32_t *keylab;
    uint8_t *keycls;
    uint32_t *rank_of_key;
    int32_t *ranklab;
    uint32_t *F[2], *E[2], *W[2], *T, *NS, *D;
    uint32_t *ctr;
    int32_t *label, *cost;
    const int8_t *lcls;
    int lcls_n, lcls_default;
    uint64_t *prof;   // [0..8] wall-clock ticks per phase kind, [16..24] phase counts
    uint64_t *bm;     // key bitmap of the step's entries
    uint32_t *wpref;  // distinct keys below each bitmap word of the step's key range
};

constexpr int NPHASE = 9;   // L1 L2 S1 S2 I1 R I2 I3 S3
// counters in the profile block: flood items and batches processed, touched-list slots claimed
constexpr int PROF_ITEMS = 240, PROF_TOUCHED = 241, PROF_BATCHES = 242, PROF_BATCHES_MAX = 243, PROF_CRIT = 244;
constexpr int PROF_BHIST = 245, NBHIST = 10;   // flood batches by ceil(log2(items)), items 1 .. 512
constexpr int PROF_CRIT_ITEMS = 255;           // items of the rounds' busiest workgroups
constexpr int PROF_FAST_HOPS = 176, PROF_FAST_TICKS = 177;   // single-wave flood hops and their wall-clock ticks

__device__ inline int label_class(const State &s, int lb)
{
    return (lb >= 0 && lb < s.lcls_n) ? s.lcls[lb] : s.lcls_default;
}

// ------------------------------------------------------------------ set-up kernels

__global__ void k_init(State s, long long n, long long nkeys)
{
    for (long long p = gtid(); p < n; p += gstride()) {
        s.word[p] = INF64;
        const bool frozen = s.label[p] != NIL;
        s.run[p] = frozen ? 0ull : INF64;   // frozen: no flood enters (see the rank-word notes)
        s.flags[p] = frozen ? FL_FROZEN : 0;
        s.qstamp[p] = 0;
        s.estamp[p] = 0;

        if (s.cost) s.cost[p] = INT32_MAX;
    }
    for (long long k = gtid(); k < nkeys; k += gstride()) s.keystamp[k] = 0;
    for (long long w = gtid(); w < nkeys / 64 + 1; w += gstride()) s.bm[w] = 0;
    if (gtid() < 256) s.prof[gtid()] = 0;
    for (long long k = gtid(); k < K_COUNT; k += gstride()) {   // one writer per counter slot
        const int i = (int)k;
        const bool inf = i == K_LEVEL || i == K_LEVEL + 1 || i == K_RSAFE || i == K_RSAFE + 1 || i == K_KW0 ||
                         i == K_KW0 + 1;
        s.ctr[i] = inf ? INF32 : 0u;
    }
}

// validate seeds, unfreeze them and flag the start of each same-label run (S order)
__global__ void k_seed_check(State s, const int32_t *__restrict__ S, int ns, long long n, uint32_t *__restrict__ runflag)
{
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int v = S[i];
        uint32_t f = 0;
        if (v < 0 || v >= n) {
            atomicOr(&s.ctr[K_ERR], 1u);
        } else {
            if (s.label[v] == NIL) atomicOr(&s.ctr[K_ERR], 2u);
            else if (label_class(s, s.label[v]) < 0) atomicOr(&s.ctr[K_ERR], 8u);
            if (atomicAdd(&s.estamp[v], 1u) != 0) atomicOr(&s.ctr[K_ERR], 4u);
            s.flags[v] = 0;
            s.run[v] = INF64;
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
            s.ctr[K_NEXTKEY] = key + 1;
            s.ctr[K_F] = (uint32_t)ns;
        }
    }
}

#include "oift_persistent.inc"
#include "oift_local.inc"

}  // namespace

namespace {

// The OIFT's device state, in context buffers reused across calls.
State alloc_state(roift_ctx &c, const Dims &d, int nadj, const LabelClasses &lc, int n_seeds, int32_t *d_label,
                  int32_t *d_cost, bool upload = true)
{
    ROIFT_REQUIRE(nadj == 7 || nadj == 27, ROIFT_EUNSUPPORTED, "adjacency must be Spheric(1.0) or Spheric(1.75)");
    ROIFT_REQUIRE(d.n > 0 && d.n < (1ll << 31), ROIFT_EINVAL, "volume must hold 1 .. 2^31-1 voxels");
    ROIFT_REQUIRE(n_seeds >= 0 && n_seeds <= d.n, ROIFT_EINVAL, "bad seed count");
    const long long n = d.n;
    State s{};
    s.word = buf<uint64_t>(c, "oift.word", n);
    s.flags = buf<uint8_t>(c, "oift.flags", n);
    s.run = buf<uint64_t>(c, "oift.run", n);

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
    // a voxel is listed once per class that reaches it: at most 2n entries, plus the
    // per-wavefront chunk slack
    s.T = buf<uint32_t>(c, "oift.T", 2 * n + (long long)MAX_BLOCKS * (BLOCK / 64) * TCHUNK);
    s.NS = buf<uint32_t>(c, "oift.NS", n);
    s.D = buf<uint32_t>(c, "oift.D", n);
    s.ctr = buf<uint32_t>(c, "oift.ctr", K_COUNT);
    s.prof = buf<uint64_t>(c, "oift.prof", 256);

    s.bm = buf<uint64_t>(c, "oift.bm", nkeys / 64 + 1);
    s.wpref = buf<uint32_t>(c, "oift.wpref", nkeys / 64 + 1);
    s.label = d_label;
    s.cost = d_cost;
    const int ncls = (int)lc.cls.size();
    int8_t *lcls = buf<int8_t>(c, "oift.lcls", ncls > 0 ? ncls : 1);
    if (ncls > 0 && upload)   // a graph capture uploads the class table before it begins
        ROIFT_HIP_CHECK(hipMemcpyAsync(lcls, lc.cls.data(), ncls, hipMemcpyHostToDevice, c.stream));
    s.lcls = lcls;
    s.lcls_n = ncls;
    s.lcls_default = lc.default_cls;
    return s;
}

}  // namespace

int run_oift_staged(roift_ctx &c, int stage, const Dims &d, int nadj, const void *d_wa, const void *d_wb,
                    int weight_dtype, const LabelClasses &lc, const int32_t *d_seeds, int n_seeds, int32_t *d_label,
                    int32_t *d_cost)
{
    ROIFT_REQUIRE(weight_dtype == ROIFT_U16 || weight_dtype == ROIFT_U32, ROIFT_EINVAL, "weight dtype must be U16 or U32");
    ROIFT_REQUIRE(n_seeds > 0, ROIFT_EINVAL, "the staged OIFT needs seeds");
    State s = alloc_state(c, d, nadj, lc, n_seeds, d_label, d_cost, stage == OIFT_PREPARE);
    if (stage == OIFT_PREPARE) {   // every buffer the launches use, allocated before a capture begins
        (void)buf<uint32_t>(c, "oift.runflag", (size_t)n_seeds);
        (void)buf<uint32_t>(c, "oift.runtot", (size_t)((n_seeds + 1023) / 1024));
        if (weight_dtype == ROIFT_U16) (void)buf<uint32_t>(c, "oift.wpk", (size_t)(nadj == 7 ? 8 : 32) * d.n);
        ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
        return 0;
    }
    const bool w16 = weight_dtype == ROIFT_U16;
    if (stage == OIFT_ENQUEUE) {
        if (nadj == 7 && w16)
            drive_enqueue<7, uint16_t>(c, s, d, (const uint16_t *)d_wa, (const uint16_t *)d_wb, d_seeds, n_seeds, false, true);
        else if (nadj == 7)
            drive_enqueue<7, uint32_t>(c, s, d, (const uint32_t *)d_wa, (const uint32_t *)d_wb, d_seeds, n_seeds, false, true);
        else if (w16)
            drive_enqueue<27, uint16_t>(c, s, d, (const uint16_t *)d_wa, (const uint16_t *)d_wb, d_seeds, n_seeds, false, true);
        else
            drive_enqueue<27, uint32_t>(c, s, d, (const uint32_t *)d_wa, (const uint32_t *)d_wb, d_seeds, n_seeds, false, true);
        return 0;
    }
    ROIFT_REQUIRE(stage == OIFT_FINISH, ROIFT_EINVAL, "bad OIFT stage");
    if (nadj == 7) return w16 ? drive_finish<7, uint16_t>(c, s, false) : drive_finish<7, uint32_t>(c, s, false);
    return w16 ? drive_finish<27, uint16_t>(c, s, false) : drive_finish<27, uint32_t>(c, s, false);
}

int run_oift(roift_ctx &c, const Dims &d, int nadj, const void *d_wa, const void *d_wb, int weight_dtype,
             const LabelClasses &lc, const int32_t *d_seeds, int n_seeds, int32_t *d_label, int32_t *d_cost)
{
    if (c.tune_oift_mode == 1) {
        // local steps for tie-free maps; on a tie the labels are restored and the
        // level-synchronous kernel runs instead
        ROIFT_REQUIRE(d.n > 0 && d.n < (1ll << 31), ROIFT_EINVAL, "volume must hold 1 .. 2^31-1 voxels");
        ROIFT_REQUIRE(n_seeds >= 0 && n_seeds <= d.n, ROIFT_EINVAL, "bad seed count");
        int32_t *lab0 = buf<int32_t>(c, "oift.label_in", d.n);
        ROIFT_HIP_CHECK(hipMemcpyAsync(lab0, d_label, 4 * d.n, hipMemcpyDeviceToDevice, c.stream));
        int r;
        if (weight_dtype == ROIFT_U16)
            r = nadj == 7 ? drive_local<7, uint16_t>(c, d, (const uint16_t *)d_wa, (const uint16_t *)d_wb, lc, d_seeds,
                                                     n_seeds, d_label, d_cost)
                          : drive_local<27, uint16_t>(c, d, (const uint16_t *)d_wa, (const uint16_t *)d_wb, lc,
                                                      d_seeds, n_seeds, d_label, d_cost);
        else
            r = nadj == 7 ? drive_local<7, uint32_t>(c, d, (const uint32_t *)d_wa, (const uint32_t *)d_wb, lc, d_seeds,
                                                     n_seeds, d_label, d_cost)
                          : drive_local<27, uint32_t>(c, d, (const uint32_t *)d_wa, (const uint32_t *)d_wb, lc,
                                                      d_seeds, n_seeds, d_label, d_cost);
        if (r >= 0) return r;
        ROIFT_HIP_CHECK(hipMemcpyAsync(d_label, lab0, 4 * d.n, hipMemcpyDeviceToDevice, c.stream));
        c.oift_local_fallbacks++;
    }
    State s = alloc_state(c, d, nadj, lc, n_seeds, d_label, d_cost);
    if (weight_dtype == ROIFT_U16) {
        if (nadj == 7) return drive<7, uint16_t>(c, s, d, (const uint16_t *)d_wa, (const uint16_t *)d_wb, d_seeds, n_seeds);
        return drive<27, uint16_t>(c, s, d, (const uint16_t *)d_wa, (const uint16_t *)d_wb, d_seeds, n_seeds);
    }
    ROIFT_REQUIRE(weight_dtype == ROIFT_U32, ROIFT_EINVAL, "weight dtype must be U16 or U32");
    if (nadj == 7) return drive<7, uint32_t>(c, s, d, (const uint32_t *)d_wa, (const uint32_t *)d_wb, d_seeds, n_seeds);
    return drive<27, uint32_t>(c, s, d, (const uint32_t *)d_wa, (const uint32_t *)d_wb, d_seeds, n_seeds);
}

namespace {

template <int NADJ, typename IT>
void launch_prep(roift_ctx &c, State &s, const Dims &d, const void *img, double fp, double fm, uint32_t *wpk,
                 int *part, uint32_t *ovf)
{
    const dim3 grid((d.xs + PX - 1) / PX, (d.ys + PY - 1) / PY, (d.zs + PZ - 1) / PZ);
    hipLaunchKernelGGL((k_prep<NADJ, IT>), grid, dim3(PX, PY), 0, c.stream, (const IT *)img, s, d, fp, fm, wpk, part,
                       ovf, 2 * d.n + 1024);
    ROIFT_HIP_CHECK(hipGetLastError());
}

}  // namespace

int run_oift_image(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, float per,
                   const LabelClasses &lc, const int32_t *d_seeds, int n_seeds, int32_t *d_label, int32_t *d_cost,
                   WeightInfo &wi)
{
    ROIFT_REQUIRE(image_dtype == ROIFT_I16 || image_dtype == ROIFT_I32, ROIFT_EINVAL, "image dtype must be I16 or I32");
    const dim3 grid((d.xs + PX - 1) / PX, (d.ys + PY - 1) / PY, (d.zs + PZ - 1) / PZ);
    const long long nblk = (long long)grid.x * grid.y * grid.z;
    // the brick flood reads the planar weights; grids beyond the launch limits take the planar path too
    const bool fused = c.tune_flood_mode != 1 && grid.y <= 65535 && grid.z <= 65535;
    if (fused) {
        State s = alloc_state(c, d, nadj, lc, n_seeds, d_label, d_cost);
        uint32_t *wpk = buf<uint32_t>(c, "oift.wpk", (size_t)(nadj == 7 ? 8 : 32) * d.n);
        int *part = buf<int>(c, "oift.prep.part", 2 * (size_t)nblk);
        int *ext = buf<int>(c, "oift.prep.ext", 4);
        uint32_t *ovf = reinterpret_cast<uint32_t *>(ext + 2);
        fill<int>(c, ext, 4, 0);
        // per_pq / 100.0 in double, as the reference's "w *= (1.0 +- per_pq/100.0)"
        const double fp = 1.0 + (double)per / 100.0, fm = 1.0 - (double)per / 100.0;
        const bool i16 = image_dtype == ROIFT_I16;
        ev_record(c, 8);
        if (nadj == 7) {
            if (i16) launch_prep<7, int16_t>(c, s, d, d_image, fp, fm, wpk, part, ovf);
            else launch_prep<7, int32_t>(c, s, d, d_image, fp, fm, wpk, part, ovf);
        } else {
            if (i16) launch_prep<27, int16_t>(c, s, d, d_image, fp, fm, wpk, part, ovf);
            else launch_prep<27, int32_t>(c, s, d, d_image, fp, fm, wpk, part, ovf);
        }
        hipLaunchKernelGGL(k_prep_fold, dim3(1), dim3(1024), 0, c.stream, (const int *)part, (int)nblk, ext);
        ROIFT_HIP_CHECK(hipGetLastError());
        ev_record(c, 9);
        int host[4];
        ROIFT_HIP_CHECK(hipMemcpyAsync(host, ext, sizeof(host), hipMemcpyDeviceToHost, c.stream));
        ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
        c.ktimes[0] = ev_ms(c, 8, 9);
        wi = WeightInfo{host[0], host[1]};
        ROIFT_REQUIRE(wi.minI >= 0, ROIFT_EINVAL, "image has negative values (shift it to >= 0 first)");
        if (host[2] == 0) {   // every weight fits 16 bits: the state is ready
            s.wpk = wpk;
            if (nadj == 7) return drive<7, uint16_t>(c, s, d, nullptr, nullptr, d_seeds, n_seeds, true);
            return drive<27, uint16_t>(c, s, d, nullptr, nullptr, d_seeds, n_seeds, true);
        }
    }
    // planar weights (32 bits, or the brick flood), then the OIFT's own initialisation
    const long long slots = (long long)((nadj - 1) / 2) * d.n;
    int wdt = ROIFT_U16;
    void *wa = buf<uint16_t>(c, "pipe.wa", slots), *wb = buf<uint16_t>(c, "pipe.wb", slots);
    try {
        wi = build_weights(c, d, nadj, d_image, image_dtype, per, wa, wb, ROIFT_U16);
    } catch (const Error &e) {
        if (e.code != ROIFT_EUNSUPPORTED) throw;
        wdt = ROIFT_U32;
        wa = buf<uint32_t>(c, "pipe.wa32", slots);
        wb = buf<uint32_t>(c, "pipe.wb32", slots);
        wi = build_weights(c, d, nadj, d_image, image_dtype, per, wa, wb, ROIFT_U32);
    }
    ROIFT_REQUIRE(wi.minI >= 0, ROIFT_EINVAL, "image has negative values (shift it to >= 0 first)");
    return run_oift(c, d, nadj, wa, wb, wdt, lc, d_seeds, n_seeds, d_label, d_cost);
}

}  // namespace roift
