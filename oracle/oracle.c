/*
 * oracle.c -- CPU restatement of roift's OIFT (f_max) + fuzzy relaxation hot path.
 *
 * THIS IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker / the
 * timed CPU baseline.  The product path (roift_amd/, libroift.so) never links
 * or calls anything in this directory.
 *
 * Every function restates the reference algorithm from its published source
 * (file:line citations are into jung0221/roift) in code written for this
 * project, and additionally emits the quantities the reference computes but
 * discards (the f_max cost map, the predecessor direction, the pop order).
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   orc_spheric          -- neighbour order vs gft_adjrel3.cpp:47-85 (closed form)
 *   orc_oift_image       -- golden fixtures produced by the compiled reference
 *                           (tests/golden), oracle/_ref when built, Kruskal
 *                           minimax closed form for per=0, brute-force optimum
 *                           on tiny graphs, forest/Bellman invariants.
 *   orc_oift_edges       -- equality with orc_oift_image on image-derived edge
 *                           weights (two different priority queues), invariants.
 *   orc_edge_weights     -- exhaustive identity check against the closed form.
 *   orc_relax            -- golden fixtures (labels), oracle/_ref (labels),
 *                           closed forms on constant images, invariants,
 *                           masked == dense equality.
 *   orc_relax_model      -- GRAD/RBF weight models: oracle/_ref labels (the
 *                           experiments compiled from their own sources),
 *                           golden fixtures, FP32 vs FP64, masked == dense.
 *   orc_gradweight_edges -- orc_oift_edges on them == the compiled
 *                           OIFT_Multi_GradWeight labels; golden fixtures.
 *   orc_relax_binary_w   -- oracle/_ref labels (ORelax_1 image / W / Wxyz
 *                           forms); golden fixtures; constant-image closed form.
 *   orc_w_edges          -- orc_oift_edges on them == the compiled OIFT(W, ...)
 *                           and OIFT(Wx, Wy, Wz, ...) labels (golden fixtures);
 *                           W = |I| differences reduce it to orc_edge_weights
 *                           with the base doubled; exhaustive factor identity.
 *   orc_set_spacing      -- the relaxations' Dpq and the blur taps under
 *                           anisotropic spacing: golden fixtures written by the
 *                           compiled reference; isotropic spacing of any size
 *                           equals unit spacing bit for bit; the taps' closed form.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <math.h>

#define ORC_NIL (-1)
#define ORC_WHITE 0
#define ORC_GRAY 1
#define ORC_BLACK 2

enum { ORC_OK = 0, ORC_EINVAL = -1, ORC_ENOMEM = -2 };

/* Voxel spacing (gft's scn->dx/dy/dz) read by the relaxations' Dpq and the Gaussian blur's
   taps; Scene32::Create's default is 1 (gft_scene32.cpp:27-29).  Set by orc_set_spacing
   before a call (the Python wrapper restores 1, 1, 1 after it). */
static float g_sp[3] = {1.0f, 1.0f, 1.0f};

int orc_set_spacing(float dx, float dy, float dz)
{
    g_sp[0] = dx; g_sp[1] = dy; g_sp[2] = dz;
    return ORC_OK;
}

static float sp_min(void)
{
    /* MIN(scn->dx, MIN(scn->dy, scn->dz)) */
    float m = g_sp[1] < g_sp[2] ? g_sp[1] : g_sp[2];
    return g_sp[0] < m ? g_sp[0] : m;
}

/* ------------------------------------------------------------------------ */
/* Adjacency: gft::AdjRel3::Spheric(r) (gft_adjrel3.cpp:47-85).              */
/* Raster order dz, dy, dx over the ball r^2 >= dx^2+dy^2+dz^2 with the      */
/* centre moved to index 0 (the others keep raster order).                   */
/* ------------------------------------------------------------------------ */
int orc_spheric(float r, int *dx, int *dy, int *dz, int cap)
{
    int r0 = (int)r, r2 = (int)(r * r);
    int n = 0, i0 = 0;
    for (int z = -r0; z <= r0; z++)
        for (int y = -r0; y <= r0; y++)
            for (int x = -r0; x <= r0; x++)
                if (x * x + y * y + z * z <= r2) {
                    if (n >= cap) return ORC_EINVAL;
                    dx[n] = x; dy[n] = y; dz[n] = z;
                    if (x == 0 && y == 0 && z == 0) i0 = n;
                    n++;
                }
    /* rotate the centre to the front */
    int cx = dx[i0], cy = dy[i0], cz = dz[i0];
    for (int i = i0; i > 0; i--) { dx[i] = dx[i - 1]; dy[i] = dy[i - 1]; dz[i] = dz[i - 1]; }
    dx[0] = cx; dy[0] = cy; dz[0] = cz;
    return n;
}

typedef struct {
    int n;          /* A->n including the centre */
    int dx[125], dy[125], dz[125];
} orc_adj;

static int adj_make(float r, orc_adj *A)
{
    A->n = orc_spheric(r, A->dx, A->dy, A->dz, 125);
    return A->n > 0 ? ORC_OK : ORC_EINVAL;
}

/* ------------------------------------------------------------------------ */
/* Oriented arc weight, OIFT_Multi (gft_ift.cpp:965-973) and the PerClass    */
/* variant (gft_ift.cpp:1055-1067): |I_p-I_q| scaled by (1 +- per_pq/100) in */
/* double and truncated toward zero by the int assignment.                   */
/* ------------------------------------------------------------------------ */
static inline int arc_weight(int Ip, int Iq, float per_pq)
{
    int w = abs(Ip - Iq);
    if (Ip > Iq)
        w = (int)((double)w * (1.0 + (double)per_pq / 100.0));
    else if (Ip < Iq)
        w = (int)((double)w * (1.0 - (double)per_pq / 100.0));
    return w;
}

static inline float label_polarity(int lb, int use_pc, float per, const float *pc, int max_label)
{
    if (!use_pc)
        return (lb == 0) ? -per : per;          /* gft_ift.cpp:968 */
    if (lb >= 0 && lb <= max_label)             /* gft_ift.cpp:1058-1062 */
        return pc[lb];
    return 0.0f;
}

/* ------------------------------------------------------------------------ */
/* Dial bucket queue with global-min FIFO pops, restating gft PQueue32's      */
/* Fast* operations (gft_pqueue32.cpp:231-245, 300-327, 353-375, 415-420,    */
/* 446-451): insert appends at the bucket tail and lowers minvalue; the pop  */
/* scans upward from minvalue and takes the bucket head.                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int *first, *last;   /* per bucket */
    int *next, *prev;    /* per element */
    char *color;
    int nbuckets, minvalue, nadded;
    const int *value;
} dialq;

static int dq_create(dialq *Q, int nbuckets, int nelems, const int *value)
{
    Q->first = (int *)malloc(sizeof(int) * (size_t)(nbuckets + 1));
    Q->last = (int *)malloc(sizeof(int) * (size_t)(nbuckets + 1));
    Q->next = (int *)malloc(sizeof(int) * (size_t)nelems);
    Q->prev = (int *)malloc(sizeof(int) * (size_t)nelems);
    Q->color = (char *)malloc((size_t)nelems);
    if (!Q->first || !Q->last || !Q->next || !Q->prev || !Q->color) return ORC_ENOMEM;
    for (int b = 0; b <= nbuckets; b++) Q->first[b] = Q->last[b] = ORC_NIL;
    for (int e = 0; e < nelems; e++) { Q->next[e] = Q->prev[e] = ORC_NIL; Q->color[e] = ORC_WHITE; }
    Q->nbuckets = nbuckets; Q->minvalue = INT_MAX; Q->nadded = 0; Q->value = value;
    return ORC_OK;
}

static void dq_destroy(dialq *Q)
{
    free(Q->first); free(Q->last); free(Q->next); free(Q->prev); free(Q->color);
}

static inline void dq_insert(dialq *Q, int e)
{
    int b = Q->value[e];
    Q->nadded++;
    if (b < Q->minvalue) Q->minvalue = b;
    if (Q->first[b] == ORC_NIL) { Q->first[b] = e; Q->prev[e] = ORC_NIL; }
    else { Q->next[Q->last[b]] = e; Q->prev[e] = Q->last[b]; }
    Q->last[b] = e;
    Q->next[e] = ORC_NIL;
    Q->color[e] = ORC_GRAY;
}

static inline void dq_remove(dialq *Q, int e)
{
    int b = Q->value[e], pv = Q->prev[e], nx = Q->next[e];
    Q->nadded--;
    if (Q->first[b] == e) {
        Q->first[b] = nx;
        if (nx == ORC_NIL) Q->last[b] = ORC_NIL; else Q->prev[nx] = ORC_NIL;
    } else {
        Q->next[pv] = nx;
        if (nx == ORC_NIL) Q->last[b] = pv; else Q->prev[nx] = pv;
    }
    Q->color[e] = ORC_BLACK;
}

static inline int dq_pop_min_fifo(dialq *Q)
{
    Q->nadded--;
    while (Q->first[Q->minvalue] == ORC_NIL) Q->minvalue++;
    int b = Q->minvalue, e = Q->first[b], nx = Q->next[e];
    if (nx == ORC_NIL) Q->first[b] = Q->last[b] = ORC_NIL;
    else { Q->first[b] = nx; Q->prev[nx] = ORC_NIL; }
    Q->color[e] = ORC_BLACK;
    return e;
}

/* ------------------------------------------------------------------------ */
/* OIFT_Multi / OIFT_Multi_PerClass (gft_ift.cpp:911-989, 991-1083).         */
/*                                                                            */
/* label:  in/out, NIL=-1 for unlabelled voxels (as the reference).           */
/* S:      gft seed array, S[0]=count, S[1..] linear indices; NULL inserts    */
/*         every labelled voxel in raster order (gft_ift.cpp:944-948).        */
/* cost:   (optional) f_max cost C(q)=max(C(pred),value(q)) assigned at pop;  */
/*         INT32_MAX for voxels never popped.                                 */
/* pred:   (optional) adjacency index i (q = pred + d_i) of the last          */
/*         successful relaxation, 0 for seeds and never-popped voxels.        */
/* value:  (optional) final last-arc value (INT_MAX / 0 as initialised).      */
/* order:  (optional) voxels in pop order; *npop receives the count.          */
/* ------------------------------------------------------------------------ */
int orc_oift_image(const int32_t *I, int xs, int ys, int zs, float adj_r,
                   int use_per_class, float per, const float *per_class, int max_label,
                   const int32_t *S, int32_t *label,
                   int32_t *cost, int8_t *pred, int32_t *value_out,
                   int32_t *order, int64_t *npop)
{
    orc_adj A;
    if (adj_make(adj_r, &A) != ORC_OK) return ORC_EINVAL;
    if (xs <= 0 || ys <= 0 || zs <= 0) return ORC_EINVAL;
    const int64_t n64 = (int64_t)xs * ys * zs;
    if (n64 > INT_MAX) return ORC_EINVAL;
    const int n = (int)n64;
    int Imax = INT_MIN, Imin = INT_MAX;
    for (int p = 0; p < n; p++) { if (I[p] > Imax) Imax = I[p]; if (I[p] < Imin) Imin = I[p]; }
    if (Imin < 0) return ORC_EINVAL; /* precondition of the bucket range (oiftrelax.cpp:263-268) */

    float max_abs = fabsf(per);
    if (use_per_class) {
        if (!per_class || max_label < 0) return ORC_EINVAL;
        max_abs = 0.0f;
        for (int c = 0; c <= max_label; c++) if (fabsf(per_class[c]) > max_abs) max_abs = fabsf(per_class[c]);
    }
    if (max_abs > 100.0f) return ORC_EINVAL;
    /* int Wmax *= double  (gft_ift.cpp:925-926, 1023-1024) */
    int Wmax = (int)((double)Imax * (1.0 + (double)max_abs / 100.0));

    if (S) {
        for (int i = 1; i <= S[0]; i++) {
            if (S[i] < 0 || S[i] >= n) return ORC_EINVAL;
            if (label[S[i]] == ORC_NIL) return ORC_EINVAL; /* would index bucket INT_MAX in the reference */
        }
    }

    int *value = (int *)malloc(sizeof(int) * (size_t)n);
    int32_t *C = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    dialq Q = {0};
    if (!value || !C || dq_create(&Q, Wmax + 2, n, value) != ORC_OK) {
        free(value); free(C); dq_destroy(&Q); return ORC_ENOMEM;
    }
    for (int p = 0; p < n; p++) {
        value[p] = (label[p] == ORC_NIL) ? INT_MAX : 0;   /* gft_ift.cpp:930-936 */
        C[p] = INT32_MAX;
        if (pred) pred[p] = 0;
    }
    if (S) {
        for (int i = 1; i <= S[0]; i++) dq_insert(&Q, S[i]);
    } else {
        for (int p = 0; p < n; p++) if (label[p] != ORC_NIL) dq_insert(&Q, p);
    }

    const int xy = xs * ys;
    int64_t k = 0;
    int level = 0;
    while (Q.nadded > 0) {
        int p = dq_pop_min_fifo(&Q);
        /* cost at pop = running flood level = max(C(pred), value(p)) */
        if (value[p] > level) level = value[p];
        C[p] = level;
        if (order) order[k] = p;
        k++;
        const int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;
        const int lp = label[p];
        const float per_pq = label_polarity(lp, use_per_class, per, per_class, max_label);
        for (int i = 1; i < A.n; i++) {
            int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
            if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;
            int q = qx + qy * xs + qz * xy;
            if (Q.color[q] == ORC_BLACK) continue;
            int w = arc_weight(I[p], I[q], per_pq);
            if (w < value[q]) {
                if (Q.color[q] == ORC_GRAY) dq_remove(&Q, q);
                value[q] = w;
                label[q] = lp;
                if (pred) pred[q] = (int8_t)i;
                dq_insert(&Q, q);
            }
        }
    }
    if (cost) memcpy(cost, C, sizeof(int32_t) * (size_t)n);
    if (value_out) memcpy(value_out, value, sizeof(int32_t) * (size_t)n);
    if (npop) *npop = k;
    free(value); free(C); dq_destroy(&Q);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Oriented edge weights for a two-polarity problem (OIFT_Multi form):       */
/* for every forward neighbour slot s (adjacency index i = (n-1)/2+1+s,      */
/* q = p + d_i) store                                                         */
/*     wa[s*N+p] = w_obj(p->q),   wb[s*N+p] = w_obj(q->p)                    */
/* where w_obj uses per_pq=+per (labels != 0) and w_bg(p->q) = w_obj(q->p).  */
/* Slots whose q lies outside the volume hold 0xFFFFFFFF.                    */
/* ------------------------------------------------------------------------ */
int orc_edge_weights(const int32_t *I, int xs, int ys, int zs, float adj_r, float per,
                     uint32_t *wa, uint32_t *wb)
{
    orc_adj A;
    if (adj_make(adj_r, &A) != ORC_OK) return ORC_EINVAL;
    const int half = (A.n - 1) / 2;
    const int64_t n = (int64_t)xs * ys * zs;
    const int xy = xs * ys;
    for (int s = 0; s < half; s++) {
        int i = half + 1 + s;
        for (int64_t p = 0; p < n; p++) {
            int px = (int)((p % xy) % xs), py = (int)((p % xy) / xs), pz = (int)(p / xy);
            int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
            if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) {
                wa[s * n + p] = 0xFFFFFFFFu; wb[s * n + p] = 0xFFFFFFFFu; continue;
            }
            int64_t q = qx + (int64_t)qy * xs + (int64_t)qz * xy;
            wa[s * n + p] = (uint32_t)arc_weight(I[p], I[q], per);
            wb[s * n + p] = (uint32_t)arc_weight(I[q], I[p], per);
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Oriented edge weights of the per-voxel weight forms OIFT(W, A, scn, ...)  */
/* (gft_ift.cpp:1142-1152) and OIFT(Wx, Wy, Wz, scn, ...) (:1244-1254): the   */
/* int base W[p] + W[q] (per-axis: the W of the arc's axis), times           */
/* (1 +- per_pq/100.0) in double by the order of I[p] and I[q], truncated;    */
/* unchanged when I[p] == I[q].  The base is symmetric, so the orc_edge_     */
/* weights layout (wa = w_obj(p->q), wb = w_obj(q->p)) holds.  The per-axis   */
/* form is defined on Spheric(1.0).                                          */
/* ------------------------------------------------------------------------ */
static inline int w_arc_weight(int base, int Ip, int Iq, float per_pq)
{
    int w = base;
    if (Ip > Iq) w = (int)((double)w * (1.0 + per_pq / 100.0));
    else if (Ip < Iq) w = (int)((double)w * (1.0 - per_pq / 100.0));
    return w;
}

int orc_w_edges(const int32_t *I, const int32_t *W, const int32_t *Wx, const int32_t *Wy, const int32_t *Wz,
                int xs, int ys, int zs, float adj_r, float per, uint32_t *wa, uint32_t *wb)
{
    orc_adj A;
    if (Wx && (!Wy || !Wz)) return ORC_EINVAL;
    if (!W && !Wx) return ORC_EINVAL;
    if (adj_make(Wx ? 1.0f : adj_r, &A) != ORC_OK) return ORC_EINVAL;
    const int half = (A.n - 1) / 2;
    const int64_t n = (int64_t)xs * ys * zs;
    const int xy = xs * ys;
    for (int s = 0; s < half; s++) {
        int i = half + 1 + s;
        const int32_t *B = W;
        if (Wx) B = A.dz[i] != 0 ? Wz : (A.dy[i] != 0 ? Wy : Wx);
        for (int64_t p = 0; p < n; p++) {
            int px = (int)((p % xy) % xs), py = (int)((p % xy) / xs), pz = (int)(p / xy);
            int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
            if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) {
                wa[s * n + p] = 0xFFFFFFFFu; wb[s * n + p] = 0xFFFFFFFFu; continue;
            }
            int64_t q = qx + (int64_t)qy * xs + (int64_t)qz * xy;
            const int base = B[p] + B[q];
            wa[s * n + p] = (uint32_t)w_arc_weight(base, I[p], I[q], per);
            wb[s * n + p] = (uint32_t)w_arc_weight(base, I[q], I[p], per);
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Oriented edge weights of OIFT_Multi_GradWeight (exp_gradient_weight.cpp:  */
/* 244-254): base wf = alpha|I_p-I_q| + (1-alpha)(G_p+G_q)*0.5 in float,      */
/* w = (int)(wf (1 +- per_pq/100)) in float, (int)wf when I_p == I_q, and a   */
/* clamp at 0.  wf is symmetric in p and q and (1 - (-per)/100) == (1 +       */
/* per/100) exactly, so w_bg(p->q) = w_obj(q->p) holds as for OIFT_Multi and  */
/* the same (wa, wb) layout serves orc_oift_edges.                            */
/* ------------------------------------------------------------------------ */
static inline int grad_arc_weight(int Ip, int Iq, int Gp, int Gq, float alpha, float per_pq)
{
    float idf = (float)abs(Ip - Iq);
    float gav = (float)(Gp + Gq) * 0.5f;
    float wf = alpha * idf + (1.0f - alpha) * gav;
    int w = (int)wf;
    if (Ip > Iq) w = (int)(wf * (1.0f + per_pq / 100.0f));
    else if (Ip < Iq) w = (int)(wf * (1.0f - per_pq / 100.0f));
    if (w < 0) w = 0;
    return w;
}

int orc_gradweight_edges(const int32_t *I, const int32_t *G, int xs, int ys, int zs, float adj_r, float alpha,
                         float per, uint32_t *wa, uint32_t *wb)
{
    orc_adj A;
    if (adj_make(adj_r, &A) != ORC_OK) return ORC_EINVAL;
    const int half = (A.n - 1) / 2;
    const int64_t n = (int64_t)xs * ys * zs;
    const int xy = xs * ys;
    for (int s = 0; s < half; s++) {
        int i = half + 1 + s;
        for (int64_t p = 0; p < n; p++) {
            int px = (int)((p % xy) % xs), py = (int)((p % xy) / xs), pz = (int)(p / xy);
            int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
            if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) {
                wa[s * n + p] = 0xFFFFFFFFu; wb[s * n + p] = 0xFFFFFFFFu; continue;
            }
            int64_t q = qx + (int64_t)qy * xs + (int64_t)qz * xy;
            wa[s * n + p] = (uint32_t)grad_arc_weight(I[p], I[q], G[p], G[q], alpha, per);
            wb[s * n + p] = (uint32_t)grad_arc_weight(I[q], I[p], G[q], G[p], alpha, per);
        }
    }
    return ORC_OK;
}

/* Global standard deviation of the image, as estimate_sigma                  */
/* (exp_gaussian_rbf_relax.cpp:194-208): integer sums, double moments, var>=1. */
float orc_estimate_sigma(const int32_t *I, int64_t n)
{
    long long sum = 0, sum2 = 0;
    for (int64_t p = 0; p < n; p++) { long long v = I[p]; sum += v; sum2 += v * v; }
    double mean = (double)sum / (double)n;
    double var = (double)sum2 / (double)n - mean * mean;
    if (var < 1.0) var = 1.0;
    return (float)sqrt(var);
}

/* ------------------------------------------------------------------------ */
/* The same Prim-style OIFT driven by explicit oriented edge weights (wa,wb  */
/* as produced above, or tie-free ranks), with a lazy binary heap keyed by   */
/* (value, insertion sequence).  Popping the minimum (value, seq) is the      */
/* Dial queue's global-min FIFO order, so this equals orc_oift_image on       */
/* image-derived weights while supporting weights up to 2^32-2.              */
/* Labels with pol_sign[label]>0 use w_obj, the others w_bg; labels outside  */
/* [0,max_label] use w_obj (as OIFT_Multi treats every label != 0).          */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t key; int32_t v; } heap_ent; /* key = value<<32 | seq_lo */

typedef struct { heap_ent *a; int64_t n, cap; } bheap;

static int bh_push(bheap *h, uint64_t key, int32_t v)
{
    if (h->n == h->cap) {
        int64_t nc = h->cap ? h->cap * 2 : 1024;
        heap_ent *na = (heap_ent *)realloc(h->a, sizeof(heap_ent) * (size_t)nc);
        if (!na) return ORC_ENOMEM;
        h->a = na; h->cap = nc;
    }
    int64_t i = h->n++;
    while (i > 0) {
        int64_t par = (i - 1) / 2;
        if (h->a[par].key <= key) break;
        h->a[i] = h->a[par]; i = par;
    }
    h->a[i].key = key; h->a[i].v = v;
    return ORC_OK;
}

static heap_ent bh_pop(bheap *h)
{
    heap_ent top = h->a[0], last = h->a[--h->n];
    int64_t i = 0;
    for (;;) {
        int64_t c = 2 * i + 1;
        if (c >= h->n) break;
        if (c + 1 < h->n && h->a[c + 1].key < h->a[c].key) c++;
        if (h->a[c].key >= last.key) break;
        h->a[i] = h->a[c]; i = c;
    }
    if (h->n > 0) h->a[i] = last;
    return top;
}

int orc_oift_edges(const uint32_t *wa, const uint32_t *wb, int xs, int ys, int zs, float adj_r,
                   const int8_t *pol_sign, int max_label,
                   const int32_t *S, int32_t *label,
                   int32_t *cost, int8_t *pred, uint32_t *value_out,
                   int32_t *order, int64_t *npop)
{
    orc_adj A;
    if (adj_make(adj_r, &A) != ORC_OK) return ORC_EINVAL;
    const int half = (A.n - 1) / 2;
    const int64_t n64 = (int64_t)xs * ys * zs;
    if (n64 > INT_MAX || n64 <= 0) return ORC_EINVAL;
    const int n = (int)n64;
    if (S) for (int i = 1; i <= S[0]; i++) {
        if (S[i] < 0 || S[i] >= n || label[S[i]] == ORC_NIL) return ORC_EINVAL;
    }
    /* 32-bit insertion sequence is enough: each successful relaxation inserts once */
    uint32_t *value = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t *seq = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    char *color = (char *)calloc((size_t)n, 1);
    bheap h = {0, 0, 0};
    if (!value || !seq || !color) { free(value); free(seq); free(color); return ORC_ENOMEM; }
    uint64_t nseq = 0;
    int rc = ORC_OK;
    for (int p = 0; p < n; p++) {
        value[p] = (label[p] == ORC_NIL) ? 0xFFFFFFFFu : 0u;
        if (cost) cost[p] = INT32_MAX;
        if (pred) pred[p] = 0;
    }
#define PUSH(q) do { seq[q] = (uint32_t)nseq; \
        if (bh_push(&h, ((uint64_t)value[q] << 32) | (uint64_t)(uint32_t)nseq, q) != ORC_OK) { rc = ORC_ENOMEM; goto done; } \
        nseq++; color[q] = ORC_GRAY; } while (0)
    if (S) { for (int i = 1; i <= S[0]; i++) PUSH(S[i]); }
    else { for (int p = 0; p < n; p++) if (label[p] != ORC_NIL) PUSH(p); }

    const int xy = xs * ys;
    int64_t k = 0;
    uint32_t level = 0;
    while (h.n > 0) {
        heap_ent e = bh_pop(&h);
        int p = e.v;
        if (color[p] == ORC_BLACK || (uint32_t)e.key != seq[p] || (uint32_t)(e.key >> 32) != value[p]) continue;
        color[p] = ORC_BLACK;
        if (value[p] > level) level = value[p];
        if (cost) cost[p] = (int32_t)level;
        if (order) order[k] = p;
        k++;
        const int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;
        const int lp = label[p];
        const int obj = (lp >= 0 && lp <= max_label && pol_sign) ? (pol_sign[lp] > 0) : (lp != 0);
        for (int i = 1; i < A.n; i++) {
            int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
            if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;
            int q = qx + qy * xs + qz * xy;
            if (color[q] == ORC_BLACK) continue;
            uint32_t w;
            if (i > half) { int s = i - half - 1; w = obj ? wa[(int64_t)s * n + p] : wb[(int64_t)s * n + p]; }
            else { int s = (A.n - i) - half - 1; w = obj ? wb[(int64_t)s * n + q] : wa[(int64_t)s * n + q]; }
            if (w < value[q]) {
                value[q] = w;
                label[q] = lp;
                if (pred) pred[q] = (int8_t)i;
                PUSH(q);
            }
        }
    }
#undef PUSH
    if (value_out) memcpy(value_out, value, sizeof(uint32_t) * (size_t)n);
    if (npop) *npop = k;
done:
    free(h.a); free(value); free(seq); free(color);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* ORelax_1_Multi / ORelax_1_Multi_PerClass (gft_ift.cpp:8655-8911,          */
/* 8913-9187), restated for float (the reference's arithmetic, operation by  */
/* operation) and double (the FP64 oracle).                                   */
/*                                                                            */
/* label:   in = OIFT labels, out = argmax labels (as the reference).         */
/* memb:    (optional) final memberships, planar [nclasses][N].               */
/* classes: (optional) class_labels (up to 64); *nclasses_out its count.      */
/* dense:   1 updates every voxel every iteration instead of the dilating    */
/*          band (bit-identical, used to test that claim).                    */
/* ------------------------------------------------------------------------ */
#define RELAX_BODY(REAL, FABS, SQRT, EXP)                                                        \
    orc_adj A;                                                                                \
    if (adj_make(adj_r, &A) != ORC_OK) return ORC_EINVAL;                                     \
    const int64_t n64 = (int64_t)xs * ys * zs;                                                \
    if (n64 > INT_MAX || n64 <= 0) return ORC_EINVAL;                                         \
    const int n = (int)n64;                                                                   \
    if (!S || S[0] <= 0) return (model == ORC_RELAX_INTENSITY) ? ORC_EINVAL : ORC_OK;       \
    int max_seed_label = 0;                                                                   \
    for (int i = 1; i <= S[0]; i++) {                                                         \
        int p = S[i];                                                                         \
        if (p >= 0 && p < n && label[p] > max_seed_label) max_seed_label = label[p];          \
    }                                                                                         \
    char *present = (char *)calloc((size_t)max_seed_label + 1, 1);                            \
    int *l2c = (int *)malloc(sizeof(int) * ((size_t)max_seed_label + 1));                    \
    int cls[64];                                                                              \
    int nc = 0;                                                                               \
    present[0] = 1;                                                                           \
    for (int i = 1; i <= S[0]; i++) {                                                         \
        int p = S[i];                                                                         \
        if (p >= 0 && p < n) { int lb = label[p]; if (lb >= 0 && lb <= max_seed_label) present[lb] = 1; } \
    }                                                                                         \
    for (int i = 0; i <= max_seed_label; i++) {                                               \
        l2c[i] = -1;                                                                          \
        if (present[i]) { if (nc >= 64) { free(present); free(l2c); return ORC_EINVAL; } l2c[i] = nc; cls[nc++] = i; } \
    }                                                                                         \
    int bg = l2c[0] < 0 ? 0 : l2c[0];                                                         \
    REAL pcm[64];                                                                             \
    float max_abs = FABS(per);                                                                \
    if (use_per_class) {                                                                      \
        max_abs = 0.0f;                                                                       \
        for (int c = 0; c < nc; c++) {                                                        \
            pcm[c] = (cls[c] <= max_label) ? (REAL)per_class[cls[c]] : (REAL)0;               \
            float a = (cls[c] <= max_label) ? fabsf(per_class[cls[c]]) : 0.0f;                \
            if (a > max_abs) max_abs = a;                                                     \
        }                                                                                     \
    }                                                                                         \
    REAL Dpq[125];                                                                            \
    for (int i = 1; i < A.n; i++) {                                                           \
        /* sqrtf(x^2 dx^2 + y^2 dy^2 + z^2 dz^2) / dmin (gft_ift.cpp:8723-8731) */            \
        REAL ax = (REAL)(A.dx[i] * A.dx[i]), ay = (REAL)(A.dy[i] * A.dy[i]), az = (REAL)(A.dz[i] * A.dz[i]); \
        Dpq[i] = SQRT(ax * (REAL)g_sp[0] * (REAL)g_sp[0] + ay * (REAL)g_sp[1] * (REAL)g_sp[1] +  \
                      az * (REAL)g_sp[2] * (REAL)g_sp[2]) / (REAL)sp_min();                    \
    }                                                                                         \
    int Imax = INT_MIN;                                                                       \
    for (int p = 0; p < n; p++) if (I[p] > Imax) Imax = I[p];                                  \
    /* ORelax_1_Multi: float Wmax = Imax; Wmax *= (1.0 + |per|/100.0) (double product);      \
       GradWeight: the same in float arithmetic (exp_gradient_weight.cpp:341-342).          */ \
    REAL Wmax = (model == ORC_RELAX_GRAD)                                                     \
        ? (REAL)((REAL)(float)Imax * ((REAL)1 + (REAL)FABS(per) / (REAL)100))                 \
        : (REAL)((double)(float)Imax * (1.0 + (double)max_abs / 100.0));                      \
    const REAL sigma2 = (REAL)sigma * (REAL)sigma;                                            \
    const size_t total = (size_t)n * (size_t)nc;                                              \
    REAL *f1 = (REAL *)calloc(total, sizeof(REAL));                                           \
    REAL *f2 = (REAL *)calloc(total, sizeof(REAL));                                           \
    int *mnodes = (int *)malloc(sizeof(int) * ((size_t)n + 1));                               \
    char *mask = (char *)calloc((size_t)n, 1);                                                \
    if (!f1 || !f2 || !mnodes || !mask) { free(f1); free(f2); free(mnodes); free(mask); free(present); free(l2c); return ORC_ENOMEM; } \
    const int xy = xs * ys;                                                                   \
    mnodes[0] = 0;                                                                            \
    for (int p = 0; p < n; p++) {                                                             \
        int lb = label[p], c = bg;                                                            \
        if (lb >= 0 && lb <= max_seed_label && l2c[lb] >= 0) c = l2c[lb];                     \
        f1[(size_t)p * nc + c] = (REAL)1;                                                     \
        int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;                               \
        for (int i = 1; i < A.n; i++) {                                                       \
            int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];                      \
            if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;     \
            int q = qx + qy * xs + qz * xy;                                                   \
            if (label[p] != label[q]) { mask[p] = 1; mnodes[++mnodes[0]] = p; break; }        \
        }                                                                                     \
    }                                                                                         \
    if (dense) { mnodes[0] = n; for (int p = 0; p < n; p++) { mnodes[p + 1] = p; mask[p] = 1; } } \
    int ninic = 1;                                                                            \
    while (ntimes > 0) {                                                                      \
        memcpy(f2, f1, total * sizeof(REAL));                                                 \
        for (int j = 1; j <= mnodes[0]; j++) {                                                \
            int p = mnodes[j];                                                                \
            int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;                           \
            size_t po = (size_t)p * nc;                                                       \
            for (int k = 0; k < nc; k++) f2[po + k] = (REAL)0;                                \
            REAL sw = (REAL)0;                                                                \
            for (int i = 1; i < A.n; i++) {                                                   \
                int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];                  \
                if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue; \
                int q = qx + qy * xs + qz * xy;                                               \
                size_t qo = (size_t)q * nc;                                                   \
                REAL w;                                                                       \
                if (model == ORC_RELAX_RBF) {                                                 \
                    REAL diff = (REAL)(I[p] - I[q]);                                          \
                    w = EXP(-(diff * diff) / sigma2) / Dpq[i];                                \
                } else {                                                                      \
                    if (model == ORC_RELAX_GRAD) {                                            \
                        REAL idf = FABS((REAL)(I[p] - I[q]));                                 \
                        REAL gav = (REAL)(G[p] + G[q]) * (REAL)0.5;                           \
                        w = (REAL)alpha * idf + ((REAL)1 - (REAL)alpha) * gav;                \
                    } else {                                                                  \
                        w = FABS((REAL)(I[p] - I[q]));                                        \
                    }                                                                         \
                    REAL s;                                                                   \
                    if (use_per_class) { s = (REAL)0; for (int k = 0; k < nc; k++) s += pcm[k] * f1[qo + k]; } \
                    else s = (REAL)per * ((REAL)2 * f1[qo + bg] - (REAL)1);                   \
                    if (I[p] > I[q]) w *= ((REAL)1 + s / (REAL)100);                           \
                    else if (I[p] < I[q]) w *= ((REAL)1 - s / (REAL)100);                      \
                    if (model == ORC_RELAX_GRAD && w < (REAL)0) w = (REAL)0;                  \
                    w = Wmax - w;                                                             \
                    if (w < (REAL)0) w = (REAL)0;                                             \
                    w = w * w; w = w * w; w = w * w;                                          \
                    w /= Dpq[i];                                                              \
                }                                                                             \
                sw += w;                                                                      \
                for (int k = 0; k < nc; k++) f2[po + k] += w * f1[qo + k];                    \
            }                                                                                 \
            if (sw > (REAL)0) { for (int k = 0; k < nc; k++) f2[po + k] /= sw; }              \
            else { for (int k = 0; k < nc; k++) f2[po + k] = f1[po + k]; }                    \
        }                                                                                     \
        REAL *t = f1; f1 = f2; f2 = t;                                                        \
        for (int i = 1; i <= S[0]; i++) {                                                     \
            int p = S[i];                                                                     \
            if (p < 0 || p >= n) continue;                                                    \
            int lb = label[p], c = bg;                                                        \
            if (lb >= 0 && lb <= max_seed_label && l2c[lb] >= 0) c = l2c[lb];                 \
            size_t po = (size_t)p * nc;                                                       \
            for (int k = 0; k < nc; k++) f1[po + k] = (REAL)0;                                \
            f1[po + c] = (REAL)1;                                                             \
        }                                                                                     \
        ntimes--;                                                                             \
        if (ntimes > 0 && !dense) {                                                           \
            int nlast = mnodes[0];                                                            \
            for (int j = ninic; j <= mnodes[0]; j++) {                                        \
                int p = mnodes[j];                                                            \
                int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;                       \
                for (int i = 1; i < A.n; i++) {                                               \
                    int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];              \
                    if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue; \
                    int q = qx + qy * xs + qz * xy;                                           \
                    if (!mask[q]) { mask[q] = 1; mnodes[++nlast] = q; }                       \
                }                                                                             \
            }                                                                                 \
            ninic = mnodes[0] + 1;                                                            \
            mnodes[0] = nlast;                                                                \
        }                                                                                     \
    }                                                                                         \
    for (int p = 0; p < n; p++) {                                                             \
        size_t po = (size_t)p * nc;                                                           \
        int best = 0;                                                                         \
        REAL bv = f1[po];                                                                     \
        for (int k = 1; k < nc; k++) if (f1[po + k] > bv) { bv = f1[po + k]; best = k; }      \
        label[p] = cls[best];                                                                 \
    }                                                                                         \
    if (memb) for (int k = 0; k < nc; k++) for (int p = 0; p < n; p++)                        \
        memb[(size_t)k * n + p] = f1[(size_t)p * nc + k];                                     \
    if (classes) for (int k = 0; k < nc; k++) classes[k] = cls[k];                            \
    if (nclasses_out) *nclasses_out = nc;                                                     \
    free(f1); free(f2); free(mnodes); free(mask); free(present); free(l2c);                   \
    return ORC_OK;

/* Weight models of the multi-class relaxation:                                */
/*  ORC_RELAX_INTENSITY  ORelax_1_Multi[_PerClass] (gft_ift.cpp:8800-8816)      */
/*  ORC_RELAX_GRAD       ORelax_1_Multi_GradWeight: base alpha|dI| +            */
/*                       (1-alpha)(G_p+G_q)/2, polarity, clamp at 0, float Wmax */
/*                       (exp_gradient_weight.cpp:407-422)                      */
/*  ORC_RELAX_RBF        ORelax_1_Multi_GaussRBF: w = exp(-d^2/sigma^2)/Dpq,    */
/*                       no polarity (exp_gaussian_rbf_relax.cpp:343-345)       */
/* GRAD and RBF leave the labels unchanged when S is empty, as the experiments  */
/* do (exp_gradient_weight.cpp:296, exp_gaussian_rbf_relax.cpp:235).            */
#define ORC_RELAX_INTENSITY 0
#define ORC_RELAX_GRAD 1
#define ORC_RELAX_RBF 2

int orc_relax_model_f32(const int32_t *I, const int32_t *G, int xs, int ys, int zs, float adj_r, int model,
                        float alpha, float sigma, int use_per_class, float per, const float *per_class,
                        int max_label, const int32_t *S, int32_t *label, int ntimes, int dense,
                        float *memb, int32_t *classes, int32_t *nclasses_out)
{
    RELAX_BODY(float, fabsf, sqrtf, expf)
}

int orc_relax_model_f64(const int32_t *I, const int32_t *G, int xs, int ys, int zs, float adj_r, int model,
                        float alpha, float sigma, int use_per_class, float per, const float *per_class,
                        int max_label, const int32_t *S, int32_t *label, int ntimes, int dense,
                        double *memb, int32_t *classes, int32_t *nclasses_out)
{
    RELAX_BODY(double, fabs, sqrt, exp)
}

int orc_relax_f32(const int32_t *I, int xs, int ys, int zs, float adj_r,
                  int use_per_class, float per, const float *per_class, int max_label,
                  const int32_t *S, int32_t *label, int ntimes, int dense,
                  float *memb, int32_t *classes, int32_t *nclasses_out)
{
    return orc_relax_model_f32(I, NULL, xs, ys, zs, adj_r, ORC_RELAX_INTENSITY, 0.0f, 1.0f, use_per_class, per,
                               per_class, max_label, S, label, ntimes, dense, memb, classes, nclasses_out);
}

int orc_relax_f64(const int32_t *I, int xs, int ys, int zs, float adj_r,
                  int use_per_class, float per, const float *per_class, int max_label,
                  const int32_t *S, int32_t *label, int ntimes, int dense,
                  double *memb, int32_t *classes, int32_t *nclasses_out)
{
    return orc_relax_model_f64(I, NULL, xs, ys, zs, adj_r, ORC_RELAX_INTENSITY, 0.0f, 1.0f, use_per_class, per,
                               per_class, max_label, S, label, ntimes, dense, memb, classes, nclasses_out);
}

/* ------------------------------------------------------------------------ */
/* ORelax_1 binary fallback (gft_ift.cpp:8498-8653): one channel            */
/* f = label/Imax, hard polarity from f(q) < 0.5, no clamp, int Wmax.        */
/* ORelax_1_Multi routes here when S is NULL or empty (gft_ift.cpp:8678).    */
/* ------------------------------------------------------------------------ */
/* Weight sources: W == NULL && Wx == NULL: the image form above (|I_p-I_q|,   */
/* int Wmax = max I (1+|per|/100), divided by Dpq).  W != NULL: ORelax_1(W,...) */
/* (gft_ift.cpp:9189-9331): base W[p]+W[q], int Wmax = 2 max W (1+|per|/100),  */
/* no Dpq.  Wx != NULL: ORelax_1(Wx,Wy,Wz,...) (gft_ift.cpp:9333-9498) on       */
/* Spheric(1.0): base W_axis[p]+W_axis[q] with the axis of the neighbour,       */
/* Wmax = 2 max(max Wx, max Wy, max Wz) (1+|per|/100), no Dpq.                  */
/* The polarity direction always follows the image.  dense: every voxel every   */
/* iteration (equal to the band for dyadic label/Imax values).                  */
int orc_relax_binary_w_f32(const int32_t *I, const int32_t *W, const int32_t *Wx, const int32_t *Wy,
                           const int32_t *Wz, int xs, int ys, int zs, float adj_r, float per,
                           const int32_t *S, int32_t *label, int ntimes, int dense, float *memb)
{
    orc_adj A;
    if (adj_make(Wx ? 1.0f : adj_r, &A) != ORC_OK) return ORC_EINVAL;
    const int64_t n64 = (int64_t)xs * ys * zs;
    if (n64 > INT_MAX || n64 <= 0) return ORC_EINVAL;
    if (Wx && (!Wy || !Wz)) return ORC_EINVAL;
    const int n = (int)n64, xy = xs * ys;
    float Dpq[125];   /* gft_ift.cpp:8515-8522 */
    for (int i = 1; i < A.n; i++)
        Dpq[i] = sqrtf((float)(A.dx[i] * A.dx[i]) * g_sp[0] * g_sp[0] + (float)(A.dy[i] * A.dy[i]) * g_sp[1] * g_sp[1] +
                       (float)(A.dz[i] * A.dz[i]) * g_sp[2] * g_sp[2]) /
                 sp_min();
    int axis[125];   /* 0 = x, 1 = y, 2 = z (the last non-zero component wins, as gft_ift.cpp:9356-9364) */
    for (int i = 1; i < A.n; i++) {
        axis[i] = 0;
        if (A.dx[i] != 0) axis[i] = 0;
        if (A.dy[i] != 0) axis[i] = 1;
        if (A.dz[i] != 0) axis[i] = 2;
    }
    const int32_t *Wa[3] = {Wx, Wy, Wz};
    int Imax = INT_MIN, Bmax = INT_MIN;
    for (int p = 0; p < n; p++) {
        if (label[p] > Imax) Imax = label[p];
        if (Wx) {
            for (int t = 0; t < 3; t++) if (Wa[t][p] > Bmax) Bmax = Wa[t][p];
        } else {
            const int b = W ? W[p] : I[p];
            if (b > Bmax) Bmax = b;
        }
    }
    if (W || Wx) Bmax *= 2;
    const int Wmax = (int)((double)Bmax * (1.0 + (double)fabsf(per) / 100.0));
    float *f1 = (float *)malloc(sizeof(float) * (size_t)n), *f2 = (float *)malloc(sizeof(float) * (size_t)n);
    int *mnodes = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    char *mask = (char *)calloc((size_t)n, 1);
    if (!f1 || !f2 || !mnodes || !mask) { free(f1); free(f2); free(mnodes); free(mask); return ORC_ENOMEM; }
    mnodes[0] = 0;
    for (int p = 0; p < n; p++) {
        f1[p] = (float)label[p] / (float)Imax;
        int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;
        for (int i = 1; i < A.n; i++) {
            int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
            if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;
            int q = qx + qy * xs + qz * xy;
            if (label[p] != label[q]) { mask[p] = 1; mnodes[++mnodes[0]] = p; break; }
        }
    }
    if (dense) { mnodes[0] = n; for (int p = 0; p < n; p++) { mnodes[p + 1] = p; mask[p] = 1; } }
    int ninic = 1;
    while (ntimes > 0) {
        memcpy(f2, f1, sizeof(float) * (size_t)n);
        for (int j = 1; j <= mnodes[0]; j++) {
            int p = mnodes[j];
            int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;
            f2[p] = 0.0f;
            float sw = 0.0f;
            for (int i = 1; i < A.n; i++) {
                int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
                if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;
                int q = qx + qy * xs + qz * xy;
                float w;
                if (Wx) w = (float)(Wa[axis[i]][p] + Wa[axis[i]][q]);
                else if (W) w = (float)(W[p] + W[q]);
                else w = (float)abs(I[p] - I[q]);
                float per_pq = (f1[q] < 0.5f) ? per : -per;
                if (I[p] > I[q]) w = (float)((double)w * (1.0 + (double)per_pq / 100.0));
                else if (I[p] < I[q]) w = (float)((double)w * (1.0 - (double)per_pq / 100.0));
                w = (float)Wmax - w;
                w = w * w; w = w * w; w = w * w;
                if (!W && !Wx) w /= Dpq[i];
                sw += w;
                f2[p] += w * f1[q];
            }
            f2[p] /= sw;
        }
        float *t = f1; f1 = f2; f2 = t;
        if (S) for (int i = 1; i <= S[0]; i++) { int p = S[i]; f1[p] = (float)label[p] / (float)Imax; }
        ntimes--;
        if (ntimes > 0 && !dense) {
            int nlast = mnodes[0];
            for (int j = ninic; j <= mnodes[0]; j++) {
                int p = mnodes[j];
                int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;
                for (int i = 1; i < A.n; i++) {
                    int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
                    if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;
                    int q = qx + qy * xs + qz * xy;
                    if (!mask[q]) { mask[q] = 1; mnodes[++nlast] = q; }
                }
            }
            ninic = mnodes[0] + 1;
            mnodes[0] = nlast;
        }
    }
    for (int p = 0; p < n; p++) label[p] = (f1[p] < 0.5f) ? 0 : 1;
    if (memb) memcpy(memb, f1, sizeof(float) * (size_t)n);
    free(f1); free(f2); free(mnodes); free(mask);
    return ORC_OK;
}

int orc_relax_binary_f32(const int32_t *I, int xs, int ys, int zs, float adj_r, float per,
                         const int32_t *S, int32_t *label, int ntimes, float *memb)
{
    return orc_relax_binary_w_f32(I, NULL, NULL, NULL, NULL, xs, ys, zs, adj_r, per, S, label, ntimes, 0, memb);
}

/* ------------------------------------------------------------------------ */
/* Differentiable two-class relaxation (no reference counterpart: the       */
/* learnable weight is defined by this project; this FP64 restatement is    */
/* its oracle).  f = background membership, init 0 on label == lab_obj and  */
/* 1 elsewhere; seeds keep their initial value.  For a non-seed voxel p:    */
/*   d = I_p - I_q,  g = exp(-d^2/sigma^2),                                  */
/*   pi = 1 + alpha*sgn(d)*(2 f_q - 1)        (soft polarity, as ORelax_1)  */
/*   w_pq = scale_q * g * pi / Dpq,                                          */
/*   f'_p = sum_q w_pq f_q / sum_q w_pq   (f'_p = f_p when the sum is 0).    */
/* Loss: binary cross-entropy of the object membership o = 1 - f against   */
/* target t, o clamped to [eps, 1-eps] (zero derivative where clamped):     */
/*   L = -(1/N) sum_p [t log o + (1-t) log(1-o)].                            */
/* Outputs L, dL/dscale (N), dL/dsigma, dL/dalpha by reverse mode over the  */
/* stored states, and the final f.                                          */
/* ------------------------------------------------------------------------ */
int orc_relax_diff_f64(const int32_t *I, int xs, int ys, int zs, float adj_r, const int32_t *S,
                       const int32_t *label, int lab_obj, int niter, const double *scale, double sigma,
                       double alpha, const double *target, double eps, double *loss, double *g_scale,
                       double *g_sigma, double *g_alpha, double *f_out)
{
    orc_adj A;
    if (adj_make(adj_r, &A) != ORC_OK) return ORC_EINVAL;
    const int64_t n64 = (int64_t)xs * ys * zs;
    if (n64 > INT_MAX || n64 <= 0 || niter < 0) return ORC_EINVAL;
    const int n = (int)n64, xy = xs * ys;
    double Dpq[125];
    for (int i = 1; i < A.n; i++)
        Dpq[i] = sqrt((double)(A.dx[i] * A.dx[i] + A.dy[i] * A.dy[i] + A.dz[i] * A.dz[i]));
    double *F = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(niter + 1));
    double *lam = (double *)calloc((size_t)n, sizeof(double)), *lam2 = (double *)calloc((size_t)n, sizeof(double));
    char *seed = (char *)calloc((size_t)n, 1);
    if (!F || !lam || !lam2 || !seed) { free(F); free(lam); free(lam2); free(seed); return ORC_ENOMEM; }
    if (S) for (int i = 1; i <= S[0]; i++) if (S[i] >= 0 && S[i] < n) seed[S[i]] = 1;
    for (int p = 0; p < n; p++) F[p] = (label[p] == lab_obj) ? 0.0 : 1.0;
    const double s2 = sigma * sigma;
#define DIFF_W(p, q, i, fq, g_, sg_, pi_, w_)                                   \
    double d_ = (double)(I[p] - I[q]);                                        \
    double g_ = exp(-(d_ * d_) / s2);                                         \
    double sg_ = (d_ > 0) ? 1.0 : ((d_ < 0) ? -1.0 : 0.0);                    \
    double pi_ = 1.0 + alpha * sg_ * (2.0 * (fq) - 1.0);                      \
    double w_ = scale[q] * g_ * pi_ / Dpq[i];
    for (int k = 0; k < niter; k++) {
        const double *f = F + (size_t)k * n;
        double *fn = F + (size_t)(k + 1) * n;
        for (int p = 0; p < n; p++) {
            if (seed[p]) { fn[p] = f[p]; continue; }
            int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;
            double a = 0.0, b = 0.0;
            for (int i = 1; i < A.n; i++) {
                int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
                if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;
                int q = qx + qy * xs + qz * xy;
                DIFF_W(p, q, i, f[q], g, sg, pi, w)
                (void)g; (void)sg; (void)pi;
                a += w * f[q];
                b += w;
            }
            fn[p] = (b != 0.0) ? a / b : f[p];
        }
    }
    const double *ff = F + (size_t)niter * n;
    double L = 0.0;
    for (int p = 0; p < n; p++) {
        double o = 1.0 - ff[p];
        int clamped = 0;
        if (o < eps) { o = eps; clamped = 1; }
        if (o > 1.0 - eps) { o = 1.0 - eps; clamped = 1; }
        L -= target[p] * log(o) + (1.0 - target[p]) * log(1.0 - o);
        /* dL/df = -dL/do, dL/do = -(1/N)(t/o - (1-t)/(1-o)) */
        lam[p] = clamped ? 0.0 : (target[p] / o - (1.0 - target[p]) / (1.0 - o)) / (double)n;
    }
    L /= (double)n;
    double gs = 0.0, ga = 0.0;
    for (int p = 0; p < n; p++) g_scale[p] = 0.0;
    for (int k = niter - 1; k >= 0; k--) {
        const double *f = F + (size_t)k * n, *fn = F + (size_t)(k + 1) * n;
        for (int p = 0; p < n; p++) lam2[p] = 0.0;
        for (int p = 0; p < n; p++) {
            if (seed[p] || lam[p] == 0.0) continue;
            int px = (p % xy) % xs, py = (p % xy) / xs, pz = p / xy;
            double b = 0.0;
            for (int i = 1; i < A.n; i++) {
                int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
                if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;
                int q = qx + qy * xs + qz * xy;
                DIFF_W(p, q, i, f[q], g, sg, pi, w)
                (void)g; (void)sg; (void)pi;
                b += w;
            }
            if (b == 0.0) { lam2[p] += lam[p]; continue; }
            const double c = lam[p] / b;
            for (int i = 1; i < A.n; i++) {
                int qx = px + A.dx[i], qy = py + A.dy[i], qz = pz + A.dz[i];
                if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;
                int q = qx + qy * xs + qz * xy;
                DIFF_W(p, q, i, f[q], g, sg, pi, w)
                const double common = f[q] - fn[p];
                const double dwdf = scale[q] * g * alpha * sg * 2.0 / Dpq[i];
                lam2[q] += c * (w + common * dwdf);
                g_scale[q] += c * common * g * pi / Dpq[i];
                gs += c * common * w * 2.0 * d_ * d_ / (s2 * sigma);
                ga += c * common * scale[q] * g * sg * (2.0 * f[q] - 1.0) / Dpq[i];
            }
        }
        double *t = lam; lam = lam2; lam2 = t;
    }
#undef DIFF_W
    *loss = L;
    *g_sigma = gs;
    *g_alpha = ga;
    if (f_out) memcpy(f_out, ff, sizeof(double) * (size_t)n);
    free(F); free(lam); free(lam2); free(seed);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* gft::Scene32::GaussianBlur (gft_filtering3.cpp:385-397): convolution     */
/* with Kernel3::SphericalGaussian(scn, R=1, s=10, f=1) normalised          */
/* (gft_filtering3.cpp:72-126): taps in Spheric(1.0) order, centre first,   */
/* val = s * expf(-f * (r2 / R2)) in float (C++'s exp(float) overload),     */
/* with r2 from the spacing-scaled displacements, normalised by their float */
/* sum.  The                                                                */
/* convolution sums (float)I[q] * val in tap order over the taps inside the */
/* volume (no renormalisation at the border) and rounds with gft's ROUND    */
/* (gft_common.h:120, the 0.5 is a double) (gft_filtering3.cpp:321-346).    */
/* ------------------------------------------------------------------------ */
int orc_gauss_kernel(float *val)
{
    orc_adj A;
    if (adj_make(1.0f, &A) != ORC_OK) return ORC_EINVAL;
    const float R = 1.0f, s = 10.0f, f = 1.0f, R2 = R * R, dmin = sp_min();
    float wt = 0.0f;
    for (int i = 0; i < A.n; i++) {
        /* displacements scaled by spacing / dmin in float (gft_filtering3.cpp:114-120) */
        const float d_x = ((float)A.dx[i] * g_sp[0]) / dmin;
        const float d_y = ((float)A.dy[i] * g_sp[1]) / dmin;
        const float d_z = ((float)A.dz[i] * g_sp[2]) / dmin;
        const float r2 = d_x * d_x + d_y * d_y + d_z * d_z;
        /* exp of a float: C++ resolves exp(float) to the float overload; at unit spacing
           (r2 in {0, 1}) the double form rounds to the same taps */
        val[i] = s * expf(-f * (r2 / R2));
    }
    for (int i = 0; i < A.n; i++) wt += val[i];
    for (int i = 0; i < A.n; i++) val[i] /= wt;
    return A.n;
}

int orc_gaussian_blur(const int32_t *I, int xs, int ys, int zs, int passes, int32_t *out)
{
    orc_adj A;
    if (adj_make(1.0f, &A) != ORC_OK) return ORC_EINVAL;
    float val[7];
    orc_gauss_kernel(val);
    const int64_t n = (int64_t)xs * ys * zs;
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!tmp) return ORC_ENOMEM;
    memcpy(out, I, sizeof(int32_t) * (size_t)n);
    for (int k = 0; k < passes; k++) {
        memcpy(tmp, out, sizeof(int32_t) * (size_t)n);
        for (int z = 0; z < zs; z++)
            for (int y = 0; y < ys; y++)
                for (int x = 0; x < xs; x++) {
                    float conv = 0.0f;
                    for (int i = 0; i < A.n; i++) {
                        int qx = x + A.dx[i], qy = y + A.dy[i], qz = z + A.dz[i];
                        if (qx < 0 || qx >= xs || qy < 0 || qy >= ys || qz < 0 || qz >= zs) continue;
                        conv += (float)tmp[qx + (int64_t)qy * xs + (int64_t)qz * xs * ys] * val[i];
                    }
                    out[x + (int64_t)y * xs + (int64_t)z * xs * ys] =
                        (conv < 0) ? (int32_t)((double)conv - 0.5) : (int32_t)((double)conv + 0.5);
                }
    }
    free(tmp);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* gft::Scene32::SobelFilter (gft_filtering3.cpp:509-552): over the linear  */
/* index range [di, N-di), di = 1 + xs + xs*ys (neighbours wrap across rows */
/* and planes as the reference's address arithmetic does; every other      */
/* voxel is 0), three integer 3x3 difference sums weighted (1 2 1; 2 4 2;   */
/* 1 2 1), each shifted right by 4, then                                    */
/*   ROUND(sqrtf((f0*f0 + f1*f1 + f2*f2) / 3.0))                            */
/* with the FP32 sum promoted to double for the division and the quotient  */
/* converted back to float for sqrtf.                                       */
/* ------------------------------------------------------------------------ */
int orc_sobel(const int32_t *I, int xs, int ys, int zs, int32_t *out)
{
    static const int n0[9] = {-1, 0, 1, -1, 0, 1, -1, 0, 1};
    static const int n1[9] = {-1, -1, -1, 0, 0, 0, 1, 1, 1};
    static const int we[9] = {1, 2, 1, 2, 4, 2, 1, 2, 1};
    const int64_t N = (int64_t)xs * ys * zs, dx = 1, dy = xs, dz = (int64_t)xs * ys, di = dx + dy + dz;
    memset(out, 0, sizeof(int32_t) * (size_t)N);
    for (int64_t i = di; i < N - di; i++) {
        int a0 = 0, a1 = 0, a2 = 0;
        for (int m = 0; m < 9; m++) {
            a0 += we[m] * (I[i + dz + dx * n0[m] + dy * n1[m]] - I[i - dz + dx * n0[m] + dy * n1[m]]);
            a1 += we[m] * (I[i + dx + dz * n0[m] + dy * n1[m]] - I[i - dx + dz * n0[m] + dy * n1[m]]);
            a2 += we[m] * (I[i + dy + dx * n0[m] + dz * n1[m]] - I[i - dy + dx * n0[m] + dz * n1[m]]);
        }
        a0 >>= 4;
        a1 >>= 4;
        a2 >>= 4;
        const float f0 = (float)a0, f1 = (float)a1, f2 = (float)a2;
        const float sum = f0 * f0 + f1 * f1 + f2 * f2;
        const float g = sqrtf((float)((double)sum / 3.0));
        out[i] = (g < 0) ? (int32_t)((double)g - 0.5) : (int32_t)((double)g + 0.5);
    }
    return ORC_OK;
}

/* The gradient normalisation of exp_gradient_weight.cpp:593-607: with Gmax = max G and
   Wmax = max of the (blurred) image, G = (int)((long long)G * Wmax / Gmax) when both are
   positive. */
int orc_grad_normalize(int32_t *G, int64_t n, int32_t image_max)
{
    int32_t gmax = INT32_MIN;
    for (int64_t p = 0; p < n; p++) if (G[p] > gmax) gmax = G[p];
    if (gmax > 0 && image_max > 0)
        for (int64_t p = 0; p < n; p++) G[p] = (int32_t)((long long)G[p] * image_max / gmax);
    return ORC_OK;
}

/* compute_bfs_edt (exp_geodesic_seeds.cpp:165-215): 6-connected BFS from the seeds of S
   whose label is > 0; unreached voxels keep INT32_MAX. */
int orc_bfs_edt(const int32_t *label, int xs, int ys, int zs, const int32_t *S, int nS, int32_t *dist)
{
    const int64_t n = (int64_t)xs * ys * zs;
    int64_t *q = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + nS + 1));
    if (!q) return ORC_ENOMEM;
    for (int64_t p = 0; p < n; p++) dist[p] = INT32_MAX;
    int64_t head = 0, tail = 0;
    for (int i = 0; i < nS; i++) {   /* a repeated seed is queued again, as in the reference */
        const int64_t p = S[i];
        if (p >= 0 && p < n && label[p] > 0) { dist[p] = 0; q[tail++] = p; }
    }
    static const int D[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    while (head < tail) {
        const int64_t p = q[head++];
        const int x = (int)(p % xs), y = (int)((p / xs) % ys), z = (int)(p / ((int64_t)xs * ys));
        for (int i = 0; i < 6; i++) {
            const int vx = x + D[i][0], vy = y + D[i][1], vz = z + D[i][2];
            if (vx < 0 || vx >= xs || vy < 0 || vy >= ys || vz < 0 || vz >= zs) continue;
            const int64_t r = vx + (int64_t)vy * xs + (int64_t)vz * xs * ys;
            if (dist[r] == INT32_MAX) { dist[r] = dist[p] + 1; q[tail++] = r; }
        }
    }
    free(q);
    return ORC_OK;
}

/* inject_boundary_seeds (oiftrelax.cpp:175-222): on the grid x, y, z = 0, stride, 2 stride,
   ... (raster order), every voxel on a volume face whose label is NIL becomes a background
   seed (label 0) appended to `out`; returns the count. */
int64_t orc_boundary_seeds(int32_t *label, int xs, int ys, int zs, int stride, int32_t *out)
{
    int64_t k = 0;
    if (stride <= 0) return 0;
    for (int z = 0; z < zs; z += stride)
        for (int y = 0; y < ys; y += stride)
            for (int x = 0; x < xs; x += stride) {
                if (!(x == 0 || x == xs - 1 || y == 0 || y == ys - 1 || z == 0 || z == zs - 1)) continue;
                const int64_t p = x + (int64_t)y * xs + (int64_t)z * xs * ys;
                if (label[p] == -1) { label[p] = 0; out[k++] = (int32_t)p; }
            }
    return k;
}

/* inject_boundary_seeds_geodesic (exp_geodesic_seeds.cpp:221-283): every face voxel (raster
   order) whose label is NIL and whose BFS distance d exceeds 2 stride becomes a background
   seed when x, y or z is a multiple of the effective stride (stride for d <= 3 stride, else
   2 stride). */
int64_t orc_boundary_seeds_geodesic(int32_t *label, int xs, int ys, int zs, int stride, const int32_t *edt,
                                    int32_t *out)
{
    int64_t k = 0;
    const int threshold = stride * 2;
    for (int z = 0; z < zs; z++)
        for (int y = 0; y < ys; y++)
            for (int x = 0; x < xs; x++) {
                if (!(x == 0 || x == xs - 1 || y == 0 || y == ys - 1 || z == 0 || z == zs - 1)) continue;
                const int64_t p = x + (int64_t)y * xs + (int64_t)z * xs * ys;
                if (label[p] != -1) continue;
                const int d = edt[p];
                if (d <= threshold) continue;
                int es;
                if (d <= stride) es = stride / 2 > 1 ? stride / 2 : 1;
                else if (d <= 3 * stride) es = stride;
                else es = stride * 2;
                if (x % es != 0 && y % es != 0 && z % es != 0) continue;
                label[p] = 0;
                out[k++] = (int32_t)p;
            }
    return k;
}

/* dilation_conditional (oiftrelax.cpp:87-169): the border = voxels of label 0 with a
   Spheric(radius) neighbour of label 1; its voxels, stably sorted by image value (ties keep
   raster order), are kept up to index (int)(float)((percentile / 100.0) * (count - 1)) and
   set to label 1.  An empty border changes nothing (the reference indexes out of range
   there).  Returns the number of voxels set. */
typedef struct { int32_t v; int32_t p; } orc_vp;
static int orc_vp_cmp(const void *a, const void *b)
{
    const orc_vp *x = (const orc_vp *)a, *y = (const orc_vp *)b;
    if (x->v != y->v) return x->v < y->v ? -1 : 1;
    return x->p < y->p ? -1 : (x->p > y->p);
}

int64_t orc_dilation_conditional(const int32_t *I, int32_t *label, int xs, int ys, int zs, float radius,
                                 float percentile)
{
    orc_adj A;
    if (adj_make(radius, &A) != ORC_OK) return -1;
    const int64_t n = (int64_t)xs * ys * zs;
    uint8_t *border = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    if (!border) return -1;
    for (int64_t p = 0; p < n; p++) {
        if (label[p] != 1) continue;
        const int x = (int)(p % xs), y = (int)((p / xs) % ys), z = (int)(p / ((int64_t)xs * ys));
        for (int i = 0; i < A.n; i++) {
            const int vx = x + A.dx[i], vy = y + A.dy[i], vz = z + A.dz[i];
            if (vx < 0 || vx >= xs || vy < 0 || vy >= ys || vz < 0 || vz >= zs) continue;
            const int64_t q = vx + (int64_t)vy * xs + (int64_t)vz * xs * ys;
            if (label[q] == 0) border[q] = 1;
        }
    }
    int64_t nb = 0;
    for (int64_t p = 0; p < n; p++) nb += border[p];
    int64_t set = 0;
    if (nb > 0) {
        orc_vp *v = (orc_vp *)malloc(sizeof(orc_vp) * (size_t)nb);
        if (!v) { free(border); return -1; }
        int64_t k = 0;
        for (int64_t p = 0; p < n; p++) if (border[p]) { v[k].v = I[p]; v[k].p = (int32_t)p; k++; }
        qsort(v, (size_t)nb, sizeof(orc_vp), orc_vp_cmp);
        const float index_percentile = (float)((percentile / 100.0) * (double)(nb - 1));
        const int64_t keep = (int64_t)(int)index_percentile + 1;
        for (int64_t i = 0; i < nb && i < keep; i++) { label[v[i].p] = 1; set++; }
        free(v);
    }
    free(border);
    return set;
}
