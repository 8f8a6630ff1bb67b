// ref_capi.cpp -- thin extern "C" wrapper (written for this project) around the
// reference gft library compiled from /root/reference/gft/src by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: it lets tests/ and the cpu_baseline leg of bench.py
// call the reference's own gft::ift::OIFT_Multi[_PerClass] and
// ORelax_1_Multi[_PerClass] (gft_ift.h:108-122, 451-465) on in-memory volumes.
// The built library lives in oracle/_ref/ (git-ignored).
#include <cstring>
#include <cstdlib>
#include <chrono>
#include "gft.h"

namespace {

// voxel spacing given to every scene made here (ref_set_spacing; gft's default is 1, 1, 1)
float g_spacing[3] = {1.0f, 1.0f, 1.0f};

gft::sScene32 *make_scene(const int *data, int xs, int ys, int zs)
{
    gft::sScene32 *s = gft::Scene32::Create(xs, ys, zs);
    std::memcpy(s->data, data, sizeof(int) * (size_t)s->n);
    s->dx = g_spacing[0];
    s->dy = g_spacing[1];
    s->dz = g_spacing[2];
    return s;
}

int *make_S(const int *S)
{
    int *c = (int *)std::calloc((size_t)S[0] + 1, sizeof(int));
    std::memcpy(c, S, sizeof(int) * ((size_t)S[0] + 1));
    return c;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

#define REF_API __attribute__((visibility("default")))

extern "C" {

// OIFT only.  label is in/out (NIL=-1 unlabelled).
REF_API int ref_oift(const int *I, int xs, int ys, int zs, float adj_r,
             int use_per_class, float per, const float *per_class, int max_label,
             const int *S, int *label)
{
    gft::sAdjRel3 *A = gft::AdjRel3::Spheric(adj_r);
    gft::sScene32 *scn = make_scene(I, xs, ys, zs);
    gft::sScene32 *lab = make_scene(label, xs, ys, zs);
    int *Sc = S ? make_S(S) : nullptr;
    if (use_per_class)
        gft::ift::OIFT_Multi_PerClass(A, scn, per_class, max_label, Sc, lab);
    else
        gft::ift::OIFT_Multi(A, scn, per, Sc, lab);
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    gft::Scene32::Destroy(&scn);
    gft::AdjRel3::Destroy(&A);
    return 0;
}

void ref_get_spacing(float *out3)
{
    out3[0] = g_spacing[0];
    out3[1] = g_spacing[1];
    out3[2] = g_spacing[2];
}

REF_API int ref_set_spacing(float dx, float dy, float dz)
{
    g_spacing[0] = dx;
    g_spacing[1] = dy;
    g_spacing[2] = dz;
    return 0;
}

// The two-label outer cut gft::ift::OIFT(A, scn, per, S, label) (gft_ift.cpp:828-908) and the
// per-voxel weight forms OIFT(W, A, scn, ...) (:1086-1167) / OIFT(Wx, Wy, Wz, scn, ...)
// (:1170-1271): W == Wx == NULL runs the first.  label is in/out.
REF_API int ref_oift_w(const int *I, const int *W, const int *Wx, const int *Wy, const int *Wz, int xs, int ys,
                       int zs, float adj_r, float per, const int *S, int *label)
{
    gft::sAdjRel3 *A = gft::AdjRel3::Spheric(adj_r);
    gft::sScene32 *scn = make_scene(I, xs, ys, zs);
    gft::sScene32 *lab = make_scene(label, xs, ys, zs);
    int *Sc = S ? make_S(S) : nullptr;
    if (Wx) {
        gft::sScene32 *sx = make_scene(Wx, xs, ys, zs), *sy = make_scene(Wy, xs, ys, zs), *sz = make_scene(Wz, xs, ys, zs);
        gft::ift::OIFT(sx, sy, sz, scn, per, Sc, lab);
        gft::Scene32::Destroy(&sx);
        gft::Scene32::Destroy(&sy);
        gft::Scene32::Destroy(&sz);
    } else if (W) {
        gft::sScene32 *sw = make_scene(W, xs, ys, zs);
        gft::ift::OIFT(sw, A, scn, per, Sc, lab);
        gft::Scene32::Destroy(&sw);
    } else {
        gft::ift::OIFT(A, scn, per, Sc, lab);
    }
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    gft::Scene32::Destroy(&scn);
    gft::AdjRel3::Destroy(&A);
    return 0;
}

// Relaxation only: label in = OIFT labels, out = relaxed labels.
REF_API int ref_relax(const int *I, int xs, int ys, int zs, float adj_r,
              int use_per_class, float per, const float *per_class, int max_label,
              const int *S, int *label, int ntimes)
{
    gft::sAdjRel3 *A = gft::AdjRel3::Spheric(adj_r);
    gft::sScene32 *scn = make_scene(I, xs, ys, zs);
    gft::sScene32 *lab = make_scene(label, xs, ys, zs);
    int *Sc = S ? make_S(S) : nullptr;
    if (use_per_class)
        gft::ift::ORelax_1_Multi_PerClass(A, scn, per_class, max_label, Sc, lab, ntimes);
    else
        gft::ift::ORelax_1_Multi(A, scn, per, Sc, lab, ntimes);
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    gft::Scene32::Destroy(&scn);
    gft::AdjRel3::Destroy(&A);
    return 0;
}

// ORelax_1 binary forms (gft_ift.cpp:8498, 9189, 9333): W == Wx == NULL -> image form
// on Spheric(adj_r); W -> per-voxel weight form; Wx/Wy/Wz -> per-axis form.
// S may be NULL only for the forms that never read it; pass S[0] = 0 for none.
REF_API int ref_relax_binary(const int *I, const int *W, const int *Wx, const int *Wy, const int *Wz, int xs, int ys,
                             int zs, float adj_r, float per, const int *S, int *label, int ntimes)
{
    gft::sAdjRel3 *A = gft::AdjRel3::Spheric(adj_r);
    gft::sScene32 *scn = make_scene(I, xs, ys, zs);
    gft::sScene32 *lab = make_scene(label, xs, ys, zs);
    int *Sc = make_S(S);
    if (Wx) {
        gft::sScene32 *sx = make_scene(Wx, xs, ys, zs), *sy = make_scene(Wy, xs, ys, zs), *sz = make_scene(Wz, xs, ys, zs);
        gft::ift::ORelax_1(sx, sy, sz, scn, per, Sc, lab, ntimes);
        gft::Scene32::Destroy(&sx);
        gft::Scene32::Destroy(&sy);
        gft::Scene32::Destroy(&sz);
    } else if (W) {
        gft::sScene32 *sw = make_scene(W, xs, ys, zs);
        gft::ift::ORelax_1(sw, A, scn, per, Sc, lab, ntimes);
        gft::Scene32::Destroy(&sw);
    } else {
        gft::ift::ORelax_1(A, scn, per, Sc, lab, ntimes);
    }
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    gft::Scene32::Destroy(&scn);
    gft::AdjRel3::Destroy(&A);
    return 0;
}

// gft::Scene32::GaussianBlur (gft_filtering3.cpp:385-397) applied `passes` times, as the CLI
// does before the OIFT (oiftrelax.cpp:368-375); and the normalized kernel it convolves with.
REF_API int ref_gaussian_blur(const int *I, int xs, int ys, int zs, int passes, int *out)
{
    gft::sScene32 *scn = make_scene(I, xs, ys, zs);
    for (int k = 0; k < passes; k++) {
        gft::sScene32 *b = gft::Scene32::GaussianBlur(scn);
        gft::Scene32::Destroy(&scn);
        scn = b;
    }
    std::memcpy(out, scn->data, sizeof(int) * (size_t)scn->n);
    gft::Scene32::Destroy(&scn);
    return 0;
}

REF_API int ref_gauss_kernel(float *vals, int cap)
{
    gft::sScene32 *scn = gft::Scene32::Create(3, 3, 3);
    scn->dx = g_spacing[0];
    scn->dy = g_spacing[1];
    scn->dz = g_spacing[2];
    gft::sKernel3 *K = gft::Kernel3::SphericalGaussian(scn, 1.0f, 10.0f, 1.0f);
    gft::sKernel3 *NK = gft::Kernel3::Normalize(K);
    const int n = NK->adj->n;
    for (int i = 0; i < n && i < cap; i++) vals[i] = NK->val[i];
    gft::Kernel3::Destroy(&NK);
    gft::Kernel3::Destroy(&K);
    gft::Scene32::Destroy(&scn);
    return n;
}

// OIFT_Multi followed by ORelax_1_Multi on the same scenes, as oiftrelax.cpp:379-401
// runs them; t_oift / t_relax receive steady_clock seconds around each call.
REF_API int ref_pipeline(const int *I, int xs, int ys, int zs, float adj_r, float per,
                 const int *S, int *label, int ntimes, double *t_oift, double *t_relax)
{
    gft::sAdjRel3 *A = gft::AdjRel3::Spheric(adj_r);
    gft::sScene32 *scn = make_scene(I, xs, ys, zs);
    gft::sScene32 *lab = make_scene(label, xs, ys, zs);
    int *Sc = make_S(S);
    double t0 = now_s();
    gft::ift::OIFT_Multi(A, scn, per, Sc, lab);
    double t1 = now_s();
    gft::ift::ORelax_1_Multi(A, scn, per, Sc, lab, ntimes);
    double t2 = now_s();
    if (t_oift) *t_oift = t1 - t0;
    if (t_relax) *t_relax = t2 - t1;
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    gft::Scene32::Destroy(&scn);
    gft::AdjRel3::Destroy(&A);
    return 0;
}

// gft::Scene32::SobelFilter.
REF_API int ref_sobel(const int *I, int xs, int ys, int zs, int *out)
{
    gft::sScene32 *scn = make_scene(I, xs, ys, zs);
    gft::sScene32 *g = gft::Scene32::SobelFilter(scn);
    std::memcpy(out, g->data, sizeof(int) * (size_t)g->n);
    gft::Scene32::Destroy(&g);
    gft::Scene32::Destroy(&scn);
    return 0;
}

// gft::Scene32::Read / Write on files.
REF_API int ref_scene_dims(const char *path, int *dims)
{
    gft::sScene32 *s = gft::Scene32::Read((char *)path);
    dims[0] = s->xsize;
    dims[1] = s->ysize;
    dims[2] = s->zsize;
    gft::Scene32::Destroy(&s);
    return 0;
}

REF_API int ref_scene_read(const char *path, int *out, long long n)
{
    gft::sScene32 *s = gft::Scene32::Read((char *)path);
    if ((long long)s->n != n) {
        gft::Scene32::Destroy(&s);
        return -1;
    }
    std::memcpy(out, s->data, sizeof(int) * (size_t)n);
    gft::Scene32::Destroy(&s);
    return 0;
}

REF_API int ref_scene_write(const char *path, const int *data, int xs, int ys, int zs)
{
    gft::sScene32 *s = make_scene(data, xs, ys, zs);
    gft::Scene32::Write(s, (char *)path);
    gft::Scene32::Destroy(&s);
    return 0;
}

} // extern "C"
