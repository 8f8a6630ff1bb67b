// relax.hip -- multi-class fuzzy relaxation (ORelax_1_Multi / _PerClass) for gfx950.
//
// The reference (gft_ift.cpp:8655-9187) updates only a band that grows one
// voxel ring per iteration.  Outside the band every neighbour carries the same
// one-hot vector, so the weighted average sum(w*1)/sum(w) equals 1 exactly and
// sum(w*0)/sum(w) equals 0: a dense update over the whole volume gives the same
// numbers and needs no mask, no mask dilation and no per-iteration memcpy.
// Seeds are reset to one-hot after each iteration, which the kernel does by
// writing the one-hot vector directly.
//
// Two classes (background + one object label) store one FP32 channel, the
// background membership; the object membership is its complement.  More
// classes store planar [nclasses][N] FP32 channels.
//
// Arithmetic follows the reference's FP32 operation order per neighbour, in
// Spheric order, with FMA contraction disabled at compile time.
//
// Bound: HBM.  Algorithmic bytes per voxel and iteration: image read
// (2 B int16) + membership read (4 B/channel) + membership write (4 B/channel).
#include <algorithm>
#include <cstdio>
#include <type_traits>

#include "ctx.hpp"

namespace roift {
namespace {

constexpr uint8_t NOSEED = 255;
constexpr int NC_MAX = 8;

// Dpq[k] = sqrtf(ax^2 dx^2 + ay^2 dy^2 + az^2 dz^2) / min(dx, dy, dz) for the Spheric
// displacements, in the reference's float operation order (gft_ift.cpp:8723-8731; the
// same in ORelax_1, :8515-8522, and in the experiments' relaxations).  Returns whether any
// Dpq differs from 1 (never for 6-adjacency at isotropic spacing).
int spheric_dpq(const Dims &d, int nadj, float *Dpq)
{
    const float dmin = std::min(d.sx, std::min(d.sy, d.sz));
    int div = 0;
    for (int k = 0; k < 27; k++) Dpq[k] = 1.0f;
    for (int k = 1; k < nadj; k++) {
        const int ax = nadj == 7 ? Adjacency<7>::dx(k) : Adjacency<27>::dx(k);
        const int ay = nadj == 7 ? Adjacency<7>::dy(k) : Adjacency<27>::dy(k);
        const int az = nadj == 7 ? Adjacency<7>::dz(k) : Adjacency<27>::dz(k);
        Dpq[k] = sqrtf((float)(ax * ax) * d.sx * d.sx + (float)(ay * ay) * d.sy * d.sy +
                       (float)(az * az) * d.sz * d.sz) /
                 dmin;
        if (Dpq[k] != 1.0f) div = 1;
    }
    return div;
}

struct RelaxArgs {
    Dims d;
    float per;          // OIFT_Multi form
    float pcm[NC_MAX];  // per-class form, mapped to compact classes
    float Wmax;
    float tA, tB;       // s/100 = tA * p_bg(q) + tB  (z-marching form)
    float Dpq[27];
    int dpq_div;        // some Dpq[k] != 1: divide (dividing by 1.0f is exact, so it is skipped otherwise)
    int use_pc;
    int nc;
    int bg;
    int cls_label[NC_MAX];
    // GradWeight model: base alpha|dI| + (1-alpha)(G_p+G_q)/2; RBF model: exp(-d^2/sigma2)
    const int32_t *G;
    float alpha;
    float sigma2;
    float inv_sigma2, inv_Dpq[27];   // RBF model: 1 / sigma2, 1 / Dpq[k]
    // RBF model, 26-adjacency march: exp(-d^2 / sigma2) by |d| for d < rbf_T (the last entry is 0,
    // as is every larger |d|); rbf_T = 0 when no table is used
    const float *rbf_tab;
    int rbf_T;
    // z-marching tiles: rows (blockDim.y) and planes per workgroup; xcd = 1 maps consecutive
    // tiles to one XCD (workgroups are dealt to the 8 XCDs round-robin), so the tiles that
    // share halo rows share an L2
    int my, mz, xcd;
};

// Weight of neighbour q in p's update under the three relaxation models
// (ORelax_1_Multi gft_ift.cpp:8800-8816; ORelax_1_Multi_GradWeight
// exp_gradient_weight.cpp:407-422; ORelax_1_Multi_GaussRBF exp_gaussian_rbf_relax.cpp:343-345),
// in the reference's FP32 operation order, except that the RBF weight multiplies by the
// reciprocals of sigma^2 and Dpq instead of dividing (one rounding each way; the two IEEE
// divisions were most of its instructions; memberships stay within 1e-6 of the FP32 oracle and
// labels equal the reference's).  s is the soft polarity of q.
template <int NADJ, int WM>
__device__ inline float model_weight(const RelaxArgs &a, int k, int Ip, int Iq, long long p, long long q, float s)
{
    float w;
    if constexpr (WM == RELAX_RBF) {   // the two divisions as products with reciprocals (within 1 ulp)
        const float diff = (float)(Ip - Iq);
        return expf(-(diff * diff) * a.inv_sigma2) * a.inv_Dpq[k];
    } else if constexpr (WM == RELAX_GRAD) {
        const float gav = (float)(a.G[p] + a.G[q]) * 0.5f;
        w = a.alpha * fabsf((float)(Ip - Iq)) + (1.0f - a.alpha) * gav;
    } else {
        w = fabsf((float)(Ip - Iq));
    }
    if (Ip > Iq) w *= (1.0f + s / 100.0f);
    else if (Ip < Iq) w *= (1.0f - s / 100.0f);
    if constexpr (WM == RELAX_GRAD) w = fmaxf(w, 0.0f);
    w = a.Wmax - w;
    if (w < 0.0f) w = 0.0f;
    w = w * w;
    w = w * w;
    w = w * w;
    if (a.dpq_div) w /= a.Dpq[k];
    return w;
}

template <int NADJ, typename IT, int WM>
__global__ __launch_bounds__(256) void k_relax2(const IT *__restrict__ img, RelaxArgs a, const float *__restrict__ fin,
                                                 float *__restrict__ fout, const uint8_t *__restrict__ seedcls,
                                                 int32_t *__restrict__ label_out)
{
    using A = Adjacency<NADJ>;
    const Dims d = a.d;
    const long long xy = (long long)d.xs * d.ys;
    for (long long p = gtid(); p < d.n; p += gstride()) {
        const uint8_t sc = seedcls[p];
        float v;
        if (sc != NOSEED) {
            v = (sc == a.bg) ? 1.0f : 0.0f;
        } else {
            const int x = (int)(p % d.xs), y = (int)((p / d.xs) % d.ys), z = (int)(p / xy);
            const int Ip = (int)img[p];
            float sw = 0.0f, acc = 0.0f;
#pragma unroll
            for (int k = 1; k < NADJ; k++) {
                const int qx = x + A::dx(k), qy = y + A::dy(k), qz = z + A::dz(k);
                if (!in_volume(d, qx, qy, qz)) continue;
                const long long q = qx + (long long)qy * d.xs + (long long)qz * xy;
                const int Iq = (int)img[q];
                const float fq = fin[q];
                // soft polarity from the neighbour's memberships (gft_ift.cpp:8800-8802, 9077-9079)
                const float s = a.use_pc ? (a.pcm[a.bg] * fq + a.pcm[1 - a.bg] * (1.0f - fq))
                                         : a.per * (2.0f * fq - 1.0f);
                const float w = model_weight<NADJ, WM>(a, k, Ip, Iq, p, q, s);
                sw += w;
                acc += w * fq;
            }
            v = (sw > 0.0f) ? acc / sw : fin[p];
        }
        fout[p] = v;
        if (label_out) label_out[p] = ((1.0f - v) > v) ? a.cls_label[1 - a.bg] : a.cls_label[a.bg];
    }
}

template <int NADJ, typename IT, int WM>
__global__ __launch_bounds__(256) void k_relaxn(const IT *__restrict__ img, RelaxArgs a, const float *__restrict__ fin,
                                                 float *__restrict__ fout, const uint8_t *__restrict__ seedcls,
                                                 int32_t *__restrict__ label_out)
{
    using A = Adjacency<NADJ>;
    const Dims d = a.d;
    const long long xy = (long long)d.xs * d.ys;
    const int nc = a.nc;
    for (long long p = gtid(); p < d.n; p += gstride()) {
        const uint8_t sc = seedcls[p];
        float v[NC_MAX];
        if (sc != NOSEED) {
            for (int c = 0; c < nc; c++) v[c] = (c == sc) ? 1.0f : 0.0f;
        } else {
            const int x = (int)(p % d.xs), y = (int)((p / d.xs) % d.ys), z = (int)(p / xy);
            const int Ip = (int)img[p];
            float sw = 0.0f;
            for (int c = 0; c < nc; c++) v[c] = 0.0f;
            for (int k = 1; k < NADJ; k++) {
                const int qx = x + A::dx(k), qy = y + A::dy(k), qz = z + A::dz(k);
                if (!in_volume(d, qx, qy, qz)) continue;
                const long long q = qx + (long long)qy * d.xs + (long long)qz * xy;
                const int Iq = (int)img[q];
                float s = 0.0f;
                if (WM != RELAX_RBF) {
                    if (a.use_pc) {
                        for (int c = 0; c < nc; c++) s += a.pcm[c] * fin[c * d.n + q];
                    } else {
                        s = a.per * (2.0f * fin[a.bg * d.n + q] - 1.0f);
                    }
                }
                const float w = model_weight<NADJ, WM>(a, k, Ip, Iq, p, q, s);
                sw += w;
                for (int c = 0; c < nc; c++) v[c] += w * fin[c * d.n + q];
            }
            if (sw > 0.0f) {
                for (int c = 0; c < nc; c++) v[c] /= sw;
            } else {
                for (int c = 0; c < nc; c++) v[c] = fin[c * d.n + p];
            }
        }
        int best = 0;
        for (int c = 0; c < nc; c++) {
            fout[c * d.n + p] = v[c];
            if (c > 0 && v[c] > v[best]) best = c;
        }
        if (label_out) label_out[p] = a.cls_label[best];
    }
}

// Seed labels and their maximum (for the class list), with seed validation.
__global__ void k_seed_labels(const int32_t *__restrict__ S, int ns, long long n, const int32_t *__restrict__ label,
                              int32_t *__restrict__ out, int *__restrict__ info)
{
    int m = INT32_MIN;
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int v = S[i];
        int lb = NIL;
        if (v >= 0 && v < n) lb = label[v];
        else atomicOr(&info[1], 1);
        out[i] = lb;
        m = max(m, lb);
    }
    m = block_reduce(m, OpMax());
    if (threadIdx.x == 0) atomicMax(&info[0], m);
}

__global__ void k_mark_present(const int32_t *__restrict__ sl, int ns, int maxl, uint8_t *__restrict__ present)
{
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int lb = sl[i];
        if (lb >= 0 && lb <= maxl) present[lb] = 1;
    }
}

// Initial memberships: one-hot of the class of the OIFT label (gft_ift.cpp:8745-8752);
// seeds' classes for the per-iteration reset (gft_ift.cpp:8839-8855).
__global__ void k_relax_init(long long n, const int32_t *__restrict__ label, const int16_t *__restrict__ l2c, int maxl,
                             int nc, int bg, float *__restrict__ f, uint8_t *__restrict__ seedcls)
{
    for (long long p = gtid(); p < n; p += gstride()) {
        const int lb = label[p];
        int c = bg;
        if (lb >= 0 && lb <= maxl && l2c[lb] >= 0) c = l2c[lb];
        if (nc == 2) {
            f[p] = (c == bg) ? 1.0f : 0.0f;
        } else {
            for (int k = 0; k < nc; k++) f[k * n + p] = (k == c) ? 1.0f : 0.0f;
        }
        seedcls[p] = NOSEED;
    }
}

__global__ void k_seed_classes(const int32_t *__restrict__ S, int ns, const int32_t *__restrict__ label,
                               const int16_t *__restrict__ l2c, int maxl, int bg, uint8_t *__restrict__ seedcls)
{
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int v = S[i];
        const int lb = label[v];
        int c = bg;
        if (lb >= 0 && lb <= maxl && l2c[lb] >= 0) c = l2c[lb];
        seedcls[v] = (uint8_t)c;
    }
}

__global__ void k_labels_from_init(long long n, const int32_t *__restrict__ label, const int16_t *__restrict__ l2c,
                                   int maxl, int bg, const int32_t *__restrict__ cls_label_dev, int32_t *__restrict__ out)
{
    for (long long p = gtid(); p < n; p += gstride()) {
        const int lb = label[p];
        int c = bg;
        if (lb >= 0 && lb <= maxl && l2c[lb] >= 0) c = l2c[lb];
        out[p] = cls_label_dev[c];
    }
}

__global__ void k_memb_out2(long long n, const float *__restrict__ f, int bg, float *__restrict__ out)
{
    for (long long p = gtid(); p < n; p += gstride()) {
        const float v = fabsf(f[p]);   // the z-marching form keeps seeds' memberships negated
        out[bg * n + p] = v;
        out[(1 - bg) * n + p] = 1.0f - v;
    }
}

// ---------------------------------------------------------------- z-marching form
// 6-adjacency, two classes: each thread owns one (x, y) column of a 64x4 tile and
// marches a chunk of z planes, keeping the z-1 / z / z+1 values in registers and
// reading its x/y neighbours (L1/L2-resident) directly.  The image is int16 (intensities
// 0..32767); a seed's membership carries its sign bit (-f, -0 for 0), so that the stencil reads
// |f| with free operand modifiers and seeds keep their membership.
constexpr int MX = 64, MY = 4, MZ = 32, MY_MAX = 8;

// Weight of neighbour q in p's update, restated for fewer operations:
//   |d| * (1 +- s/100) = |d| + d*t  with d = I_p - I_q and t = s/100 = tA*f_q + tB
// (the polarity's sign is the sign of d).  Rounding differs from the reference's
// FP32 order by a few ulp; results stay within 1e-5 of the FP64 oracle.  wk is Wmax, or -inf
// for a neighbour outside the volume (its weight is then +0).
__device__ inline float relax_w8(float wk, float d, float fq, const RelaxArgs &a)
{
    const float t = fmaf(a.tA, fq, a.tB);
    float w = wk - fmaf(d, t, fabsf(d));
    w = fmaxf(w, 0.0f);
    w = w * w;
    w = w * w;
    return w * w;
}

// Batched form: one launch relaxes a batch of volumes stored one after another
// (flat, unpadded).  Block b takes tile tiles[b] of volume tiles[b].vol, whose
// offset, size and Wmax come from vols[].
struct MarchTile {
    int vol, x0, y0, z0;
};
struct MarchVol {
    long long off;
    int xs, ys, zs;
    float wmax;   // the volume's float Wmax (set on the device from its image maximum)
    int lab;      // the object label (set on the device from the seeds)
};

// DPQ: anisotropic spacing, each neighbour's weight divided by its axis's Dpq (1 = -z, 2 = -y,
// 3 = -x, 4 = +x, 5 = +y, 6 = +z in Spheric order) after the eighth power, as the reference does.
template <bool BATCH, bool DPQ>
__global__ __launch_bounds__(MX *MY_MAX) void k_relax2_march(const int16_t *__restrict__ img, RelaxArgs a,
                                                          const MarchTile *__restrict__ tiles,
                                                          const MarchVol *__restrict__ vols,
                                                          const float *__restrict__ fin, float *__restrict__ fout,
                                                          int32_t *__restrict__ label_out)
{
    int xs, ys, zs, x, y, z0;
    long long base = 0;
    if constexpr (BATCH) {
        const MarchTile t = tiles[blockIdx.x];
        const MarchVol mv = vols[t.vol];
        xs = mv.xs;
        ys = mv.ys;
        zs = mv.zs;
        base = mv.off;
        a.Wmax = mv.wmax;
        a.cls_label[1 - a.bg] = mv.lab;
        x = t.x0 + (int)threadIdx.x;
        y = t.y0 + (int)threadIdx.y;
        z0 = t.z0;
    } else {
        xs = a.d.xs;
        ys = a.d.ys;
        zs = a.d.zs;
        const unsigned ntx = (unsigned)(xs + MX - 1) / MX, nty = (unsigned)(ys + a.my - 1) / a.my;
        const unsigned ntiles = ntx * nty * ((unsigned)(zs + a.mz - 1) / a.mz);
        unsigned tile = blockIdx.x;
        if (a.xcd) tile = (blockIdx.x % 8u) * ((ntiles + 7u) / 8u) + blockIdx.x / 8u;
        if (tile >= ntiles) return;   // no block-level synchronisation below
        x = (int)(tile % ntx) * MX + (int)threadIdx.x;
        y = (int)((tile / ntx) % nty) * a.my + (int)threadIdx.y;
        z0 = (int)(tile / (ntx * nty)) * a.mz;
    }
    const int z1 = min(zs, z0 + (BATCH ? MZ : a.mz));
    if (x >= xs || y >= ys || z0 >= zs) return;   // no block-level synchronisation below
    const uint32_t xy = (uint32_t)(xs * ys), uxs = (uint32_t)xs;
    // Plane z's arrays from uniform (scalar) base pointers that advance one plane per step, and
    // loop-invariant 32-bit in-plane offsets: a load takes the scalar base and the vector offset
    // as they are, with no per-step vector address arithmetic.
    const size_t zoff = (size_t)base + (size_t)z0 * xy;
    const int16_t *iz = img + zoff;
    const float *fz = fin + zoff;
    float *oz = fout + zoff;
    int32_t *lz = label_out ? label_out + zoff : nullptr;
    // loads and stores at 32-bit byte offsets from a base pointer
    auto LI = [](const int16_t *b, uint32_t o) -> uint32_t { return *(const uint16_t *)((const char *)b + (size_t)o); };
    auto LF = [](const float *b, uint32_t o) -> float { return *(const float *)((const char *)b + (size_t)o); };
    // 2^23 + intensity as a float: one OR instead of a conversion; the difference of two such
    // values is the exact integer difference (both lie in [2^23, 2^24))
    auto BI = [](uint32_t r) -> float { return __uint_as_float(r | 0x4B000000u); };
    const int lab_bg = a.cls_label[a.bg], lab_ob = a.cls_label[1 - a.bg];
    // The in-plane neighbours are loaded unconditionally, so that all of a voxel's loads are in
    // flight together (loads under the bounds tests would each wait for their own data): an
    // absent neighbour reads the voxel itself, which is in the volume, and its Wmax is -inf, so
    // its weight max(-inf, 0)^8 is +0.  Adding a +0 weight and a +0 product leaves both sums
    // bit-identical, since the weights and memberships are non-negative.
    const uint32_t o0 = (uint32_t)x + (uint32_t)y * uxs;
    const uint32_t oym = y > 0 ? o0 - uxs : o0, oxm = x > 0 ? o0 - 1u : o0, oxp = x < xs - 1 ? o0 + 1u : o0,
                   oyp = y < ys - 1 ? o0 + uxs : o0;
    // byte offsets of the column and its in-plane neighbours: int16 image (h*) and float (b*)
    uint32_t h0 = o0 << 1, hym = oym << 1, hxm = oxm << 1, hxp = oxp << 1, hyp = oyp << 1;
    uint32_t b0 = o0 << 2, bym = oym << 2, bxm = oxm << 2, bxp = oxp << 2, byp = oyp << 2;
    const float ninf = -__builtin_huge_valf();
    const float wym = y > 0 ? a.Wmax : ninf, wxm = x > 0 ? a.Wmax : ninf, wxp = x < xs - 1 ? a.Wmax : ninf,
                wyp = y < ys - 1 ? a.Wmax : ninf;
    bool hm = z0 > 0;
    float Bm = 0.0f, fm = 0.0f;
    if (hm) {
        Bm = BI(LI(iz - xy, h0));
        fm = LF(fz - xy, b0);
    }
    uint32_t rc = LI(iz, h0);
    float fc = LF(fz, b0);
    for (int z = z0; z < z1; z++, iz += xy, fz += xy, oz += xy) {
        // re-issued per step (no instruction), so that each load pairs a 32-bit offset with the
        // scalar base instead of taking a 64-bit address hoisted out of the loop
        asm volatile("" : "+v"(h0), "+v"(hym), "+v"(hxm), "+v"(hxp), "+v"(hyp));
        asm volatile("" : "+v"(b0), "+v"(bym), "+v"(bxm), "+v"(bxp), "+v"(byp));
        // plane z+1, or plane z again at the last plane (uniform choice; its weight is skipped)
        const bool hp = z + 1 < zs;
        const uint32_t rn = LI(hp ? iz + xy : iz, h0);
        const float fn = LF(hp ? fz + xy : fz, b0);
        const uint32_t rym = LI(iz, hym), rxm = LI(iz, hxm), rxp = LI(iz, hxp), ryp = LI(iz, hyp);
        const float fym = LF(fz, bym), fxm = LF(fz, bxm), fxp = LF(fz, bxp), fyp = LF(fz, byp);
        const float Bc = BI(rc);
        float v = fc;
        if (!(__float_as_uint(fc) >> 31)) {
            // neighbours in Spheric order: -z, -y, -x, +x, +y, +z (the reference's summation order)
            float sw = 0.0f, acc = 0.0f, w;
            if (hm) {
                w = relax_w8(a.Wmax, Bc - Bm, fabsf(fm), a);
                if constexpr (DPQ) w /= a.Dpq[1];
                sw += w;
                acc = fmaf(w, fabsf(fm), acc);
            }
            w = relax_w8(wym, Bc - BI(rym), fabsf(fym), a);
            if constexpr (DPQ) w /= a.Dpq[2];
            sw += w;
            acc = fmaf(w, fabsf(fym), acc);
            w = relax_w8(wxm, Bc - BI(rxm), fabsf(fxm), a);
            if constexpr (DPQ) w /= a.Dpq[3];
            sw += w;
            acc = fmaf(w, fabsf(fxm), acc);
            w = relax_w8(wxp, Bc - BI(rxp), fabsf(fxp), a);
            if constexpr (DPQ) w /= a.Dpq[4];
            sw += w;
            acc = fmaf(w, fabsf(fxp), acc);
            w = relax_w8(wyp, Bc - BI(ryp), fabsf(fyp), a);
            if constexpr (DPQ) w /= a.Dpq[5];
            sw += w;
            acc = fmaf(w, fabsf(fyp), acc);
            if (hp) {
                w = relax_w8(a.Wmax, Bc - BI(rn), fabsf(fn), a);
                if constexpr (DPQ) w /= a.Dpq[6];
                sw += w;
                acc = fmaf(w, fabsf(fn), acc);
            }
            v = (sw > 0.0f) ? acc / sw : fc;
        }
        *(float *)((char *)oz + (size_t)b0) = v;
        if (lz) {
            const float av = fabsf(v);
            *(int32_t *)((char *)lz + (size_t)b0) = ((1.0f - av) > av) ? lab_ob : lab_bg;
            lz += xy;
        }
        hm = true;
        Bm = Bc;
        fm = fc;
        rc = rn;
        fc = fn;
    }
}

// 26-adjacency, two classes, any weight model: each thread owns one (x, y) column of a 64 x my
// tile and marches a chunk of z planes, keeping the 3 x 3 neighbourhood of planes z-1, z and z+1
// in registers, so a voxel costs 9 image and 9 membership loads (plane z+1) instead of 26 + 26.
// A neighbour outside the volume reads the voxel's own column and its weight is zeroed: a +0
// weight and a +0 product leave both sums bit-identical to skipping it (weights and memberships
// are >= 0), so the result equals k_relax2's, summed in the reference's Spheric order (the
// raster order of the 3 x 3 x 3 block without its centre).
// RBF weights from a table: the model's exp(-d^2 / sigma2) depends on the integer |d| alone, so
// for intensities in [0, 32767] one LDS read by |d| (from v_sad_u16) replaces the exponential,
// bit-identical to it; |d| >= rbf_T - 1 reads the table's last entry, 0, as the exponential of
// anything below -105 is.  Absent columns add 2^30 to |d|, so they read the 0 as well.
constexpr int RELAX_RBF_TAB = 3;
constexpr int RBF_TMAX = 4096;

__global__ void k_rbf_table(float *__restrict__ tab, int T, float inv_sigma2)
{
    for (int d = (int)gtid(); d < T; d += (int)gstride()) {
        const float diff = (float)d;   // model_weight's expression
        tab[d] = expf(-(diff * diff) * inv_sigma2);
    }
}

template <typename IT, int WM>
__global__ __launch_bounds__(MX *MY_MAX) void k_relax2_march26(const IT *__restrict__ img, RelaxArgs a,
                                                            const float *__restrict__ fin, float *__restrict__ fout,
                                                            const uint8_t *__restrict__ seedcls,
                                                            int32_t *__restrict__ label_out)
{
    constexpr bool TAB = WM == RELAX_RBF_TAB;
    const int xs = a.d.xs, ys = a.d.ys, zs = a.d.zs;
    const unsigned ntx = (unsigned)(xs + MX - 1) / MX, nty = (unsigned)(ys + a.my - 1) / a.my;
    const unsigned ntiles = ntx * nty * ((unsigned)(zs + a.mz - 1) / a.mz);
    unsigned tile = blockIdx.x;
    if (a.xcd) tile = (blockIdx.x % 8u) * ((ntiles + 7u) / 8u) + blockIdx.x / 8u;
    if (tile >= ntiles) return;   // no block-level synchronisation below
    const int x = (int)(tile % ntx) * MX + (int)threadIdx.x;
    const int y = (int)((tile / ntx) % nty) * a.my + (int)threadIdx.y;
    const int z0 = (int)(tile / (ntx * nty)) * a.mz;
    const int z1 = min(zs, z0 + a.mz);
    __shared__ float tab[TAB ? RBF_TMAX : 1];
    if constexpr (TAB) {   // every thread of the workgroup is still here
        for (int i = (int)(threadIdx.x + threadIdx.y * MX); i < a.rbf_T; i += MX * (int)blockDim.y) tab[i] = a.rbf_tab[i];
        __syncthreads();
    }
    if (x >= xs || y >= ys || z0 >= zs) return;   // no block-level synchronisation below
    const uint32_t uxs = (uint32_t)xs, xy = (uint32_t)xs * (uint32_t)ys;
    // the 3 x 3 columns around (x, y): c = (dy + 1) * 3 + (dx + 1); absent ones read column 4
    int32_t off[9];
    uint32_t msk[9];
#pragma unroll
    for (int c = 0; c < 9; c++) {
        const int dx = c % 3 - 1, dy = c / 3 - 1;
        const bool ok = (unsigned)(x + dx) < (unsigned)xs && (unsigned)(y + dy) < (unsigned)ys;
        off[c] = ok ? dx + dy * xs : 0;
        msk[c] = ok ? ~0u : 0u;
    }
    int I[3][9] = {};
    float F[3][9] = {};
    const uint32_t pc = (uint32_t)x + (uint32_t)y * uxs;   // the column's in-plane offset
    // Byte offsets of the 3 x 3 columns (image and memberships), loop-invariant; plane z's arrays
    // come from uniform (scalar) base pointers, so a load takes the scalar base and the vector
    // offset with no per-step address arithmetic.
    uint32_t hof[9], bof[9];
#pragma unroll
    for (int c = 0; c < 9; c++) {
        hof[c] = (uint32_t)((int)pc + off[c]) * (uint32_t)sizeof(IT);
        bof[c] = (uint32_t)((int)pc + off[c]) * 4u;
    }
    uint32_t pb = pc * 4u, p1 = pc;
    // the planes' register slots rotate with the unrolled steps below, not by copies
    auto load_plane = [&](auto S, int z) {
        constexpr int sl = decltype(S)::value;
        const IT *ib = img + (size_t)z * xy;
        const float *fb = fin + (size_t)z * xy;
#pragma unroll
        for (int c = 0; c < 9; c++) {
            I[sl][c] = (int)*(const IT *)((const char *)ib + (size_t)hof[c]);
            F[sl][c] = *(const float *)((const char *)fb + (size_t)bof[c]);
        }
    };
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    using S2 = std::integral_constant<int, 2>;
    if (z0 > 0) load_plane(S0{}, z0 - 1);
    load_plane(S1{}, z0);
    const int lab_bg = a.cls_label[a.bg], lab_ob = a.cls_label[1 - a.bg];
    // one voxel: planes z-1, z and z+1 sit in slots A, B and C
    auto step = [&](auto SA, auto SB, auto SC, int z) {
        constexpr int A = decltype(SA)::value, B = decltype(SB)::value;
        // re-issued per step (no instruction), so that each access pairs a 32-bit offset with
        // the scalar base instead of taking a 64-bit address hoisted out of the loop
        asm volatile("" : "+v"(hof[0]), "+v"(hof[1]), "+v"(hof[2]), "+v"(hof[3]), "+v"(hof[4]), "+v"(hof[5]),
                     "+v"(hof[6]), "+v"(hof[7]), "+v"(hof[8]));
        asm volatile("" : "+v"(bof[0]), "+v"(bof[1]), "+v"(bof[2]), "+v"(bof[3]), "+v"(bof[4]), "+v"(bof[5]),
                     "+v"(bof[6]), "+v"(bof[7]), "+v"(bof[8]), "+v"(pb), "+v"(p1));
        const bool hm = z > 0, hp = z + 1 < zs;
        if (hp) load_plane(SC, z + 1);
        const size_t zo = (size_t)z * xy;
        const uint32_t p = pc + (uint32_t)z * xy;
        const uint8_t sc = *(seedcls + zo + p1);
        float v;
        if (sc != NOSEED) {
            v = (sc == a.bg) ? 1.0f : 0.0f;
        } else if constexpr (TAB) {
            // the same sums in the same order; an absent plane's terms (+0) are skipped
            const uint32_t lim = (uint32_t)a.rbf_T - 1u, Ip = (uint32_t)I[B][4];
            float sw = 0.0f, acc = 0.0f;
#pragma unroll
            for (int pl = 0; pl < 3; pl++) {
                if ((pl == 0 && !hm) || (pl == 2 && !hp)) continue;   // uniform
                const int sl = pl == 0 ? A : (pl == 1 ? B : decltype(SC)::value);
#pragma unroll
                for (int c = 0; c < 9; c++) {
                    if (pl == 1 && c == 4) continue;
                    const int j = pl * 9 + c, k = j < 13 ? j + 1 : j;   // Spheric index
                    const uint32_t ad = __builtin_amdgcn_sad_u16(Ip, (uint32_t)I[sl][c], msk[c] ? 0u : 0x40000000u);
                    const float w = tab[min(ad, lim)] * a.inv_Dpq[k];
                    sw += w;
                    acc += w * F[sl][c];
                }
            }
            v = (sw > 0.0f) ? acc / sw : F[B][4];
        } else {
            const int Ip = I[B][4];
            float sw = 0.0f, acc = 0.0f;
            int k = 1;
#pragma unroll
            for (int pl = 0; pl < 3; pl++) {
                const uint32_t pm = (pl == 0 ? hm : (pl == 2 ? hp : true)) ? ~0u : 0u;
                const int sl = pl == 0 ? A : (pl == 1 ? B : decltype(SC)::value);
#pragma unroll
                for (int c = 0; c < 9; c++) {
                    if (pl == 1 && c == 4) continue;
                    const uint32_t m = msk[c] & pm;
                    const float fq = F[sl][c];
                    const long long q = m ? (long long)((int)p + off[c] + (pl - 1) * (int)xy) : (long long)p;
                    // soft polarity from the neighbour's memberships (gft_ift.cpp:8800-8802, 9077-9079)
                    const float s = a.use_pc ? (a.pcm[a.bg] * fq + a.pcm[1 - a.bg] * (1.0f - fq))
                                             : a.per * (2.0f * fq - 1.0f);
                    float w = model_weight<27, WM>(a, k, Ip, I[sl][c], p, q, s);
                    w = __uint_as_float(__float_as_uint(w) & m);
                    sw += w;
                    acc += w * __uint_as_float(__float_as_uint(fq) & m);
                    k++;
                }
            }
            v = (sw > 0.0f) ? acc / sw : F[B][4];
        }
        *(float *)((char *)(fout + zo) + (size_t)pb) = v;
        if (label_out) *(int32_t *)((char *)(label_out + zo) + (size_t)pb) = ((1.0f - v) > v) ? lab_ob : lab_bg;
    };
    for (int z = z0; z < z1;) {
        step(S0{}, S1{}, S2{}, z);
        if (++z >= z1) break;
        step(S1{}, S2{}, S0{}, z);
        if (++z >= z1) break;
        step(S2{}, S0{}, S1{}, z);
        ++z;
    }
}

// int16 copy of the image for the z-marching form (intensities 0..32767), and the seeds' sign
// bit on their initial membership
template <typename IT>
__global__ void k_pack_image(const IT *__restrict__ img, const uint8_t *__restrict__ seedcls, long long n,
                             int16_t *__restrict__ out, float *__restrict__ f)
{
    for (long long p = gtid(); p < n; p += gstride()) {
        out[p] = (int16_t)img[p];
        if (seedcls[p] != NOSEED) f[p] = -f[p];
    }
}

template <int NADJ, typename IT, int WM>
void launch_iter_m(roift_ctx &c, const void *img, const RelaxArgs &a, const float *fin, float *fout,
                   const uint8_t *seedcls, int32_t *label_out)
{
    const unsigned g = grid_for(a.d.n, 256, 1 << 16);
    if (a.nc == 2)
        hipLaunchKernelGGL((k_relax2<NADJ, IT, WM>), dim3(g), dim3(256), 0, c.stream, (const IT *)img, a, fin, fout,
                           seedcls, label_out);
    else
        hipLaunchKernelGGL((k_relaxn<NADJ, IT, WM>), dim3(g), dim3(256), 0, c.stream, (const IT *)img, a, fin, fout,
                           seedcls, label_out);
    ROIFT_HIP_CHECK(hipGetLastError());
}

template <int NADJ, typename IT>
void launch_iter(roift_ctx &c, const void *img, const RelaxArgs &a, int model, const float *fin, float *fout,
                 const uint8_t *seedcls, int32_t *label_out)
{
    if (model == RELAX_GRAD) launch_iter_m<NADJ, IT, RELAX_GRAD>(c, img, a, fin, fout, seedcls, label_out);
    else if (model == RELAX_RBF) launch_iter_m<NADJ, IT, RELAX_RBF>(c, img, a, fin, fout, seedcls, label_out);
    else launch_iter_m<NADJ, IT, RELAX_INTENSITY>(c, img, a, fin, fout, seedcls, label_out);
}

template <typename IT>
void launch_march26(roift_ctx &c, const void *img, const RelaxArgs &a, int model, dim3 grid, const float *fin,
                    float *fout, const uint8_t *seedcls, int32_t *label_out)
{
    const dim3 blk(MX, a.my);
    if (model == RELAX_GRAD)
        hipLaunchKernelGGL((k_relax2_march26<IT, RELAX_GRAD>), grid, blk, 0, c.stream, (const IT *)img, a, fin, fout,
                           seedcls, label_out);
    else if (model == RELAX_RBF && a.rbf_T > 0)
        hipLaunchKernelGGL((k_relax2_march26<IT, RELAX_RBF_TAB>), grid, blk, 0, c.stream, (const IT *)img, a, fin,
                           fout, seedcls, label_out);
    else if (model == RELAX_RBF)
        hipLaunchKernelGGL((k_relax2_march26<IT, RELAX_RBF>), grid, blk, 0, c.stream, (const IT *)img, a, fin, fout,
                           seedcls, label_out);
    else
        hipLaunchKernelGGL((k_relax2_march26<IT, RELAX_INTENSITY>), grid, blk, 0, c.stream, (const IT *)img, a, fin,
                           fout, seedcls, label_out);
    ROIFT_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- ORelax_1 (binary)
// Single channel f = label/Imax; hard polarity from f(q) < 0.5; the arc factor is applied
// in double and no clamp precedes the eighth power (gft_ift.cpp:8570-8596).  Seeds keep
// their initial value (the reference resets them to label/Imax after every iteration).
// Weight base: the image (|I_p-I_q|, divided by Dpq), a per-voxel weight W (W_p+W_q,
// gft_ift.cpp:9189-9331) or the per-axis weights of the neighbour's axis (gft_ift.cpp:9333-9498).
enum : int { BIN_IMAGE = 0, BIN_W = 1, BIN_WXYZ = 2 };

struct BinArgs {
    Dims d;
    float per;
    float Wmax;   // (float) of the reference's int Wmax
    float Dpq[27];
    int dpq_div;  // as RelaxArgs::dpq_div
    const int32_t *W;
    const int32_t *Wa[3];   // x, y, z
};

template <int NADJ, typename IT, int BM>
__global__ __launch_bounds__(256) void k_relax_bin(const IT *__restrict__ img, BinArgs a, const float *__restrict__ fin,
                                                    float *__restrict__ fout, const uint8_t *__restrict__ seed,
                                                    int32_t *__restrict__ label_out)
{
    using A = Adjacency<NADJ>;
    const Dims d = a.d;
    const long long xy = (long long)d.xs * d.ys;
    for (long long p = gtid(); p < d.n; p += gstride()) {
        float v;
        if (seed[p]) {
            v = fin[p];
        } else {
            const int x = (int)(p % d.xs), y = (int)((p / d.xs) % d.ys), z = (int)(p / xy);
            const int Ip = (int)img[p];
            float sw = 0.0f, acc = 0.0f;
#pragma unroll
            for (int k = 1; k < NADJ; k++) {
                const int qx = x + A::dx(k), qy = y + A::dy(k), qz = z + A::dz(k);
                if (!in_volume(d, qx, qy, qz)) continue;
                const long long q = qx + (long long)qy * d.xs + (long long)qz * xy;
                const int Iq = (int)img[q];
                const float fq = fin[q];
                float w;
                if constexpr (BM == BIN_WXYZ) {
                    const int t = A::dz(k) != 0 ? 2 : (A::dy(k) != 0 ? 1 : 0);
                    w = (float)(a.Wa[t][p] + a.Wa[t][q]);
                } else if constexpr (BM == BIN_W) {
                    w = (float)(a.W[p] + a.W[q]);
                } else {
                    w = (float)abs(Ip - Iq);
                }
                const float per_pq = (fq < 0.5f) ? a.per : -a.per;
                if (Ip > Iq) w = (float)((double)w * (1.0 + (double)per_pq / 100.0));
                else if (Ip < Iq) w = (float)((double)w * (1.0 - (double)per_pq / 100.0));
                w = a.Wmax - w;
                w = w * w;
                w = w * w;
                w = w * w;
                if (BM == BIN_IMAGE && a.dpq_div) w /= a.Dpq[k];
                sw += w;
                acc += w * fq;
            }
            v = acc / sw;
        }
        fout[p] = v;
        if (label_out) label_out[p] = (v < 0.5f) ? 0 : 1;
    }
}

__global__ void k_label_max(const int32_t *__restrict__ label, long long n, int *__restrict__ out)
{
    int m = INT32_MIN;
    for (long long p = gtid(); p < n; p += gstride()) m = max(m, label[p]);
    m = block_reduce(m, OpMax());
    if (threadIdx.x == 0) atomicMax(out, m);
}

__global__ void k_bin_init(const int32_t *__restrict__ label, long long n, float imax, float *__restrict__ f,
                           uint8_t *__restrict__ seed)
{
    for (long long p = gtid(); p < n; p += gstride()) {
        f[p] = (float)label[p] / imax;
        seed[p] = 0;
    }
}

__global__ void k_bin_seeds(const int32_t *__restrict__ S, int ns, long long n, uint8_t *__restrict__ seed,
                            int *__restrict__ bad)
{
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int v = S[i];
        if (v >= 0 && v < n) seed[v] = 1;
        else atomicOr(bad, 1);
    }
}

__global__ void k_bin_labels(const float *__restrict__ f, long long n, int32_t *__restrict__ label)
{
    for (long long p = gtid(); p < n; p += gstride()) label[p] = (f[p] < 0.5f) ? 0 : 1;
}

template <int NADJ, typename IT>
void launch_bin(roift_ctx &c, int bm, const void *img, const BinArgs &a, const float *fin, float *fout,
                const uint8_t *seed, int32_t *lo)
{
    const unsigned g = grid_for(a.d.n, 256, 1 << 16);
    if (bm == BIN_WXYZ)
        hipLaunchKernelGGL((k_relax_bin<NADJ, IT, BIN_WXYZ>), dim3(g), dim3(256), 0, c.stream, (const IT *)img, a, fin,
                           fout, seed, lo);
    else if (bm == BIN_W)
        hipLaunchKernelGGL((k_relax_bin<NADJ, IT, BIN_W>), dim3(g), dim3(256), 0, c.stream, (const IT *)img, a, fin,
                           fout, seed, lo);
    else
        hipLaunchKernelGGL((k_relax_bin<NADJ, IT, BIN_IMAGE>), dim3(g), dim3(256), 0, c.stream, (const IT *)img, a,
                           fin, fout, seed, lo);
    ROIFT_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- batched two-class relaxation
__global__ void k_batch_finit(const int32_t *__restrict__ label, long long n, const int *__restrict__ lab_obj,
                              float *__restrict__ f)
{
    const int lo = *lab_obj;
    for (long long p = gtid(); p < n; p += gstride()) f[p] = (label[p] == lo) ? 0.0f : 1.0f;
}

// Each volume's float Wmax = (float)((double)(float)max I * (1.0 + |per| / 100.0)), as
// gft_ift.cpp:8733-8734 rounds it, from the device extrema (ext[2v] = max of volume v), and the
// batch's object label.
__global__ void k_batch_vols(MarchVol *__restrict__ vols, int nvol, const int *__restrict__ ext,
                             const int *__restrict__ lab_obj, float per)
{
    for (int v = (int)gtid(); v < nvol; v += (int)gstride()) {
        vols[v].wmax = (float)((double)(float)ext[2 * v] * (1.0 + (double)fabsf(per) / 100.0));
        vols[v].lab = *lab_obj;
    }
}

template <typename IT>
__global__ void k_batch_pack(const IT *__restrict__ img, long long n, int16_t *__restrict__ out)
{
    for (long long p = gtid(); p < n; p += gstride()) out[p] = (int16_t)((int)img[p] & 0x7FFF);
}

// the seeds' sign bit on their initial membership (setting a bit: listing a seed twice is harmless)
__global__ void k_batch_seedbits(const int32_t *__restrict__ S, int ns, long long n, float *__restrict__ f)
{
    for (long long i = gtid(); i < ns; i += gstride()) {
        const int v = S[i];
        if (v < 0 || v >= n) continue;   // an invalid seed (reported by the caller's checks)
        f[v] = __uint_as_float(__float_as_uint(f[v]) | 0x80000000u);
    }
}

}  // namespace

int run_relax(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, const RelaxSpec &rs,
              const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter, float *d_memb, int32_t *nclasses_out,
              const WeightInfo &wi, const RelaxModel &rm)
{
    const int32_t maxI = wi.maxI;
    ROIFT_REQUIRE(rm.model == RELAX_INTENSITY || !rs.use_per_class, ROIFT_EINVAL,
                  "the gradient-weight and RBF relaxations take a single polarity");
    ROIFT_REQUIRE(rm.model != RELAX_GRAD || rm.G != nullptr, ROIFT_EINVAL, "gradient-weight relaxation needs G");
    ROIFT_REQUIRE(nadj == 7 || nadj == 27, ROIFT_EUNSUPPORTED, "adjacency must be Spheric(1.0) or Spheric(1.75)");
    ROIFT_REQUIRE(n_seeds > 0, ROIFT_EINVAL, "relaxation needs seeds (the S==NULL form is ORelax_1)");
    ROIFT_REQUIRE(niter >= 0, ROIFT_EINVAL, "negative iteration count");
    const long long n = d.n;
    hipStream_t st = c.stream;
    // --- classes: seed labels present, plus 0 (gft_ift.cpp:8684-8720)
    int32_t *sl = buf<int32_t>(c, "relax.seedlab", n_seeds);
    int *info = buf<int>(c, "relax.info", 2);
    const int info0[2] = {0, 0};
    ROIFT_HIP_CHECK(hipMemcpyAsync(info, info0, sizeof(info0), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_seed_labels, dim3(grid_for(n_seeds)), dim3(256), 0, st, d_seeds, n_seeds, n, d_label, sl, info);
    ROIFT_HIP_CHECK(hipGetLastError());
    int hinfo[2];
    ROIFT_HIP_CHECK(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    ROIFT_REQUIRE(hinfo[1] == 0, ROIFT_EINVAL, "seed index out of range");
    const int maxl = std::max(0, hinfo[0]);
    ROIFT_REQUIRE(maxl < (1 << 20), ROIFT_EUNSUPPORTED, "seed labels above 2^20");
    uint8_t *present = buf<uint8_t>(c, "relax.present", (size_t)maxl + 1);
    fill<uint8_t>(c, present, maxl + 1, 0);
    hipLaunchKernelGGL(k_mark_present, dim3(grid_for(n_seeds)), dim3(256), 0, st, sl, n_seeds, maxl, present);
    std::vector<uint8_t> hp((size_t)maxl + 1);
    ROIFT_HIP_CHECK(hipMemcpyAsync(hp.data(), present, hp.size(), hipMemcpyDeviceToHost, st));
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    hp[0] = 1;
    std::vector<int> cls;
    std::vector<int16_t> l2c((size_t)maxl + 1, -1);
    for (int l = 0; l <= maxl; l++)
        if (hp[l]) {
            l2c[l] = (int16_t)cls.size();
            cls.push_back(l);
        }
    const int nc = (int)cls.size();
    ROIFT_REQUIRE(nc <= NC_MAX, ROIFT_EUNSUPPORTED, "more than 8 relaxation classes");
    if (nclasses_out) *nclasses_out = nc;
    const int bg = l2c[0];  // label 0 is always class 0

    RelaxArgs a{};
    a.d = d;
    a.per = rs.per;
    a.use_pc = rs.use_per_class;
    a.nc = nc;
    a.bg = bg;
    float max_abs = std::fabs(rs.per);
    if (rs.use_per_class) {
        max_abs = 0.f;
        for (int k = 0; k < nc; k++) {
            const int L = cls[k];
            a.pcm[k] = (L <= rs.max_label) ? rs.per_class[L] : 0.0f;
            max_abs = std::max(max_abs, std::fabs(a.pcm[k]));
        }
    }
    // float Wmax = (float)Imax; Wmax *= (1.0 + |per|/100.0)   (gft_ift.cpp:8733-8734);
    // the GradWeight relaxation computes the same in float (exp_gradient_weight.cpp:341-342)
    a.Wmax = (float)((double)(float)maxI * (1.0 + (double)max_abs / 100.0));
    if (rm.model == RELAX_GRAD) a.Wmax = (float)maxI * (1.0f + std::fabs(rs.per) / 100.0f);
    a.G = rm.G;
    a.alpha = rm.alpha;
    a.sigma2 = rm.sigma * rm.sigma;
    a.dpq_div = spheric_dpq(d, nadj, a.Dpq);
    a.inv_sigma2 = 1.0f / a.sigma2;
    for (int k = 0; k < 27; k++) a.inv_Dpq[k] = a.Dpq[k] != 0.0f ? 1.0f / a.Dpq[k] : 0.0f;
    for (int k = 0; k < nc; k++) a.cls_label[k] = cls[k];
    if (nc == 2) {
        if (rs.use_per_class) {
            a.tA = (float)(((double)a.pcm[bg] - (double)a.pcm[1 - bg]) / 100.0);
            a.tB = (float)((double)a.pcm[1 - bg] / 100.0);
        } else {
            a.tA = (float)(2.0 * (double)rs.per / 100.0);
            a.tB = (float)(-(double)rs.per / 100.0);
        }
    }

    int16_t *dl2c = buf<int16_t>(c, "relax.l2c", (size_t)maxl + 1);
    ROIFT_HIP_CHECK(hipMemcpyAsync(dl2c, l2c.data(), l2c.size() * sizeof(int16_t), hipMemcpyHostToDevice, st));
    const long long chans = (nc == 2) ? 1 : nc;
    float *f0 = buf<float>(c, "relax.f0", (size_t)(chans * n));
    float *f1 = buf<float>(c, "relax.f1", (size_t)(chans * n));
    uint8_t *seedcls = buf<uint8_t>(c, "relax.seedcls", n);
    const unsigned G = grid_for(n, 256, 8192);
    hipLaunchKernelGGL(k_relax_init, dim3(G), dim3(256), 0, st, n, d_label, dl2c, maxl, nc, bg, f0, seedcls);
    hipLaunchKernelGGL(k_seed_classes, dim3(grid_for(n_seeds)), dim3(256), 0, st, d_seeds, n_seeds, d_label, dl2c, maxl,
                       bg, seedcls);
    ROIFT_HIP_CHECK(hipGetLastError());
    if (niter == 0) {
        int32_t *cl = buf<int32_t>(c, "relax.clslab", nc);
        ROIFT_HIP_CHECK(hipMemcpyAsync(cl, cls.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_labels_from_init, dim3(G), dim3(256), 0, st, n, d_label, dl2c, maxl, bg, cl, d_label);
    }
    float *fin = f0, *fout = f1;
    const bool i16 = image_dtype == ROIFT_I16;
    // z-marching form: 6-adjacency, two classes, intensities in [0, 32767] (bit 15 = seed)
    const bool march = rm.model == RELAX_INTENSITY && nadj == 7 && nc == 2 && wi.minI >= 0 && wi.maxI <= 0x7FFF &&
                       n < (1ll << 31) && (long long)d.xs * d.ys < (1ll << 30) && niter > 0;
    // any weight model; 32-bit byte offsets within a plane
    const bool march26 = nadj == 27 && nc == 2 && n < (1ll << 31) && (long long)d.xs * d.ys < (1ll << 30);
    if (march26 && rm.model == RELAX_RBF && wi.minI >= 0 && wi.maxI <= 0x7FFF) {
        // the first |d| whose exponent is below -105 (float products, as the kernels form them):
        // it and every larger |d| give exactly 0
        int T = 0;
        for (int dd = 0; dd < RBF_TMAX; dd++) {
            const float fd = (float)dd;
            if ((fd * fd) * a.inv_sigma2 >= 105.0f) {
                T = dd + 1;
                break;
            }
        }
        if (T > 0) {
            float *tab = buf<float>(c, "relax.rbftab", RBF_TMAX);
            hipLaunchKernelGGL(k_rbf_table, dim3(grid_for(T)), dim3(256), 0, st, tab, T, a.inv_sigma2);
            ROIFT_HIP_CHECK(hipGetLastError());
            a.rbf_tab = tab;
            a.rbf_T = T;
        }
    }
    int16_t *pimg = nullptr;
    if (march) {
        pimg = buf<int16_t>(c, "relax.pimg", n);
        if (i16)
            hipLaunchKernelGGL(k_pack_image<int16_t>, dim3(G), dim3(256), 0, st, (const int16_t *)d_image, seedcls, n, pimg,
                               f0);
        else
            hipLaunchKernelGGL(k_pack_image<int32_t>, dim3(G), dim3(256), 0, st, (const int32_t *)d_image, seedcls, n, pimg,
                               f0);
        ROIFT_HIP_CHECK(hipGetLastError());
    }
    // tile shape: ROIFT_PARAM_RELAX_TILE = rows | planes << 8 | xcd << 16 (0: 4 rows, 32 planes, XCD-grouped)
    a.my = MY;
    a.mz = MZ;
    a.xcd = 1;
    if (c.tune_relax_tile > 0) {
        a.my = std::min(MY_MAX, std::max(1, c.tune_relax_tile & 0xFF));
        a.mz = std::max(1, (c.tune_relax_tile >> 8) & 0xFF);
        a.xcd = (c.tune_relax_tile >> 16) & 1;
    }
    const unsigned ntiles = (unsigned)((d.xs + MX - 1) / MX) * (unsigned)((d.ys + a.my - 1) / a.my) *
                            (unsigned)((d.zs + a.mz - 1) / a.mz);
    const dim3 mgrid(a.xcd ? 8u * ((ntiles + 7u) / 8u) : ntiles);
    ev_record(c, 12);
    for (int it = 0; it < niter; it++) {
        int32_t *lo = (it == niter - 1) ? d_label : nullptr;
        if (march) {
            if (a.dpq_div)
                hipLaunchKernelGGL((k_relax2_march<false, true>), mgrid, dim3(MX, a.my), 0, st, pimg, a, nullptr,
                                   nullptr, fin, fout, lo);
            else
                hipLaunchKernelGGL((k_relax2_march<false, false>), mgrid, dim3(MX, a.my), 0, st, pimg, a, nullptr,
                                   nullptr, fin, fout, lo);
            ROIFT_HIP_CHECK(hipGetLastError());
        } else if (nadj == 7) {
            if (i16) launch_iter<7, int16_t>(c, d_image, a, rm.model, fin, fout, seedcls, lo);
            else launch_iter<7, int32_t>(c, d_image, a, rm.model, fin, fout, seedcls, lo);
        } else if (march26) {
            if (i16) launch_march26<int16_t>(c, d_image, a, rm.model, mgrid, fin, fout, seedcls, lo);
            else launch_march26<int32_t>(c, d_image, a, rm.model, mgrid, fin, fout, seedcls, lo);
        } else {
            if (i16) launch_iter<27, int16_t>(c, d_image, a, rm.model, fin, fout, seedcls, lo);
            else launch_iter<27, int32_t>(c, d_image, a, rm.model, fin, fout, seedcls, lo);
        }
        std::swap(fin, fout);
    }
    ev_record(c, 13);
    if (d_memb) {
        if (nc == 2)
            hipLaunchKernelGGL(k_memb_out2, dim3(G), dim3(256), 0, st, n, fin, bg, d_memb);
        else
            ROIFT_HIP_CHECK(hipMemcpyAsync(d_memb, fin, sizeof(float) * nc * n, hipMemcpyDeviceToDevice, st));
        ROIFT_HIP_CHECK(hipGetLastError());
    }
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    c.ktimes[2] = ev_ms(c, 12, 13);
    c.ktimes[3] = niter;
    return 0;
}

}  // namespace roift

namespace roift {

int run_relax_binary(roift_ctx &c, const Dims &d, int nadj, const void *d_image, int image_dtype, float per,
                     const int32_t *W, const int32_t *Wx, const int32_t *Wy, const int32_t *Wz,
                     const int32_t *d_seeds, int n_seeds, int32_t *d_label, int niter, float *d_memb)
{
    ROIFT_REQUIRE(nadj == 7 || nadj == 27, ROIFT_EUNSUPPORTED, "adjacency must be Spheric(1.0) or Spheric(1.75)");
    ROIFT_REQUIRE(niter >= 0 && n_seeds >= 0, ROIFT_EINVAL, "negative iteration or seed count");
    const bool wxyz = Wx != nullptr;
    ROIFT_REQUIRE(!wxyz || (Wy && Wz), ROIFT_EINVAL, "per-axis weights need Wx, Wy and Wz");
    ROIFT_REQUIRE(!wxyz || !W, ROIFT_EINVAL, "give either W or Wx/Wy/Wz");
    // the per-axis form always runs on Spheric(1.0) (gft_ift.cpp:9355)
    if (wxyz) nadj = 7;
    const long long n = d.n;
    hipStream_t st = c.stream;
    const int bm = wxyz ? BIN_WXYZ : (W ? BIN_W : BIN_IMAGE);
    // Wmax (int in the reference): max I, or 2 max W, or 2 max(max Wx, max Wy, max Wz), times (1+|per|/100)
    int base_max;
    if (bm == BIN_IMAGE) {
        base_max = image_extrema(c, d, d_image, image_dtype).maxI;
    } else if (bm == BIN_W) {
        base_max = image_extrema(c, d, W, ROIFT_I32).maxI * 2;
    } else {
        base_max = std::max(image_extrema(c, d, Wx, ROIFT_I32).maxI,
                            std::max(image_extrema(c, d, Wy, ROIFT_I32).maxI, image_extrema(c, d, Wz, ROIFT_I32).maxI)) *
                   2;
    }
    BinArgs a{};
    a.d = d;
    a.per = per;
    a.Wmax = (float)(int)((double)base_max * (1.0 + (double)std::fabs(per) / 100.0));
    a.dpq_div = spheric_dpq(d, nadj, a.Dpq);
    a.W = W;
    a.Wa[0] = Wx;
    a.Wa[1] = Wy;
    a.Wa[2] = Wz;
    // Imax = max label (gft_ift.cpp:8525)
    int *info = buf<int>(c, "bin.info", 2);
    const int info0[2] = {INT32_MIN, 0};
    ROIFT_HIP_CHECK(hipMemcpyAsync(info, info0, sizeof(info0), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_label_max, dim3(grid_for(n, 256, 1024)), dim3(256), 0, st, d_label, n, info);
    ROIFT_HIP_CHECK(hipGetLastError());
    int hinfo[2];
    ROIFT_HIP_CHECK(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    float *f0 = buf<float>(c, "relax.f0", n);
    float *f1 = buf<float>(c, "relax.f1", n);
    uint8_t *seed = buf<uint8_t>(c, "relax.seedcls", n);
    const unsigned G = grid_for(n, 256, 8192);
    hipLaunchKernelGGL(k_bin_init, dim3(G), dim3(256), 0, st, d_label, n, (float)hinfo[0], f0, seed);
    if (n_seeds > 0)
        hipLaunchKernelGGL(k_bin_seeds, dim3(grid_for(n_seeds)), dim3(256), 0, st, d_seeds, n_seeds, n, seed, info + 1);
    ROIFT_HIP_CHECK(hipGetLastError());
    const bool i16 = image_dtype == ROIFT_I16;
    float *fin = f0, *fout = f1;
    ev_record(c, 12);
    for (int it = 0; it < niter; it++) {
        int32_t *lo = (it == niter - 1) ? d_label : nullptr;
        if (nadj == 7) {
            if (i16) launch_bin<7, int16_t>(c, bm, d_image, a, fin, fout, seed, lo);
            else launch_bin<7, int32_t>(c, bm, d_image, a, fin, fout, seed, lo);
        } else {
            if (i16) launch_bin<27, int16_t>(c, bm, d_image, a, fin, fout, seed, lo);
            else launch_bin<27, int32_t>(c, bm, d_image, a, fin, fout, seed, lo);
        }
        std::swap(fin, fout);
    }
    ev_record(c, 13);
    if (niter == 0) {
        hipLaunchKernelGGL(k_bin_labels, dim3(G), dim3(256), 0, st, fin, n, d_label);
        ROIFT_HIP_CHECK(hipGetLastError());
    }
    if (d_memb) ROIFT_HIP_CHECK(hipMemcpyAsync(d_memb, fin, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
    ROIFT_HIP_CHECK(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
    ROIFT_HIP_CHECK(hipStreamSynchronize(st));
    ROIFT_REQUIRE(hinfo[1] == 0, ROIFT_EINVAL, "seed index out of range");
    c.ktimes[2] = ev_ms(c, 12, 13);
    c.ktimes[3] = niter;
    return 0;
}

}  // namespace roift

namespace roift {

int run_relax_batch(roift_ctx &c, const std::vector<BatchVolume> &vols, const void *d_image, int image_dtype,
                    float per, const int *d_lab_obj, const int *d_ext, const int32_t *d_seeds, int n_seeds,
                    int32_t *d_label, int niter, int mode)
{
    ROIFT_REQUIRE(!vols.empty() && niter >= 0, ROIFT_EINVAL, "empty batch or negative iteration count");
    const bool capture = mode == RELAX_BATCH_CAPTURE;
    hipStream_t st = c.stream;
    long long n = 0;
    std::vector<MarchVol> mv;
    std::vector<MarchTile> tiles;
    for (size_t v = 0; v < vols.size(); v++) {
        const BatchVolume &b = vols[v];
        ROIFT_REQUIRE(b.off == n, ROIFT_EINVAL, "batch volumes must be stored back to back");
        mv.push_back(MarchVol{b.off, b.xs, b.ys, b.zs, 0.0f, 0});   // wmax and lab: k_batch_vols
        for (int z = 0; z < b.zs; z += MZ)
            for (int y = 0; y < b.ys; y += MY)
                for (int x = 0; x < b.xs; x += MX) tiles.push_back(MarchTile{(int)v, x, y, z});
        n += (long long)b.xs * b.ys * b.zs;
    }
    ROIFT_REQUIRE(n < (1ll << 31), ROIFT_EINVAL, "batch larger than 2^31-1 voxels");
    MarchVol *dvol = buf<MarchVol>(c, "batch.vols", mv.size());
    MarchTile *dtile = buf<MarchTile>(c, "batch.tiles", tiles.size());
    if (!capture) {   // a graph's tables are uploaded outside it (they depend on the shapes only)
        ROIFT_HIP_CHECK(hipMemcpyAsync(dvol, mv.data(), mv.size() * sizeof(MarchVol), hipMemcpyHostToDevice, st));
        ROIFT_HIP_CHECK(hipMemcpyAsync(dtile, tiles.data(), tiles.size() * sizeof(MarchTile), hipMemcpyHostToDevice, st));
    }
    float *f0 = buf<float>(c, "relax.f0", n);
    float *f1 = buf<float>(c, "relax.f1", n);
    int16_t *pimg = buf<int16_t>(c, "relax.pimg", n);
    if (mode == RELAX_BATCH_TABLES) return 0;
    RelaxArgs a{};
    a.d = Dims{0, 0, 0, n};
    a.per = per;
    a.nc = 2;
    a.bg = 0;   // classes {0, lab_obj} in increasing label order; the object label comes from MarchVol
    a.cls_label[0] = 0;
    a.tA = (float)(2.0 * (double)per / 100.0);
    a.tB = (float)(-(double)per / 100.0);
    const unsigned G = grid_for(n, 256, 8192);
    hipLaunchKernelGGL(k_batch_vols, dim3(1), dim3(256), 0, st, dvol, (int)mv.size(), d_ext, d_lab_obj, per);
    hipLaunchKernelGGL(k_batch_finit, dim3(G), dim3(256), 0, st, d_label, n, d_lab_obj, f0);
    if (image_dtype == ROIFT_I16)
        hipLaunchKernelGGL(k_batch_pack<int16_t>, dim3(G), dim3(256), 0, st, (const int16_t *)d_image, n, pimg);
    else
        hipLaunchKernelGGL(k_batch_pack<int32_t>, dim3(G), dim3(256), 0, st, (const int32_t *)d_image, n, pimg);
    if (n_seeds > 0)
        hipLaunchKernelGGL(k_batch_seedbits, dim3(grid_for(n_seeds)), dim3(256), 0, st, d_seeds, n_seeds, n, f0);
    ROIFT_HIP_CHECK(hipGetLastError());
    const unsigned nt = (unsigned)tiles.size();
    if (!capture) ev_record(c, 12);
    float *fin = f0, *fout = f1;
    for (int it = 0; it < niter; it++) {
        int32_t *lo = (it == niter - 1) ? d_label : nullptr;
        hipLaunchKernelGGL((k_relax2_march<true, false>), dim3(nt), dim3(MX, MY), 0, st, pimg, a, dtile, dvol, fin,
                           fout, lo);
        std::swap(fin, fout);
    }
    ROIFT_HIP_CHECK(hipGetLastError());
    if (!capture) {
        ev_record(c, 13);
        ROIFT_HIP_CHECK(hipStreamSynchronize(st));
        c.ktimes[2] = ev_ms(c, 12, 13);
        c.ktimes[3] = niter;
    }
    return 0;
}

}  // namespace roift
