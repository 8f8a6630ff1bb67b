// driver.cpp -- test driver for libroift_gft.so, the gft-signature drop-in (test infrastructure).
//
// A caller of the reference links its own code against gft; here the caller is this file:
// it builds gft's records (sScene32, sAdjRel3) the way gft lays them out (gft_scene32.h:20-27,
// gft_adjrel3.h:10-23), defines the one gft constructor the drop-in resolves from the caller's
// link (Scene32::Create, used by GaussianBlur for its result), and calls the drop-in through
// the reference's own C++ signatures (include/roift_gft.hpp).  Each drv_* entry takes plain
// arrays so the Python tests can call it through ctypes in their own process.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/roift_gft.hpp"

namespace gft {
namespace Scene32 {
// Test-side stand-in for gft's constructor: zeroed int data, unit spacing.  Exported, so the
// drop-in's reference to gft::Scene32::Create binds to it as it would to libgft's.
__attribute__((visibility("default"))) sScene32 *Create(int xsize, int ysize, int zsize)
{
    sScene32 *s = (sScene32 *)std::calloc(1, sizeof(sScene32));
    s->xsize = xsize;
    s->ysize = ysize;
    s->zsize = zsize;
    s->n = xsize * ysize * zsize;
    s->dx = s->dy = s->dz = 1.0f;
    s->data = (int *)std::calloc((size_t)s->n, sizeof(int));
    return s;
}
}  // namespace Scene32
}  // namespace gft

namespace {

struct Scene {
    gft::sScene32 s{};
    explicit Scene(int *data, int xs, int ys, int zs, const float *sp = nullptr)
    {
        s.data = data;
        s.xsize = xs;
        s.ysize = ys;
        s.zsize = zs;
        s.n = xs * ys * zs;
        s.dx = sp ? sp[0] : 1.0f;
        s.dy = sp ? sp[1] : 1.0f;
        s.dz = sp ? sp[2] : 1.0f;
    }
};

// gft::AdjRel3::Spheric(r): raster order over the ball with the centre first (gft_adjrel3.cpp:47-85)
struct Adj {
    std::vector<gft::AdjRel3::sDisplacement3> d;
    gft::sAdjRel3 a{};
    explicit Adj(float r)
    {
        const int r0 = (int)r, r2 = (int)(r * r);
        gft::AdjRel3::sDisplacement3 c{};
        d.push_back(c);
        for (int z = -r0; z <= r0; z++)
            for (int y = -r0; y <= r0; y++)
                for (int x = -r0; x <= r0; x++)
                    if ((x || y || z) && x * x + y * y + z * z <= r2) {
                        gft::AdjRel3::sDisplacement3 v{};
                        v.axis.x = x;
                        v.axis.y = y;
                        v.axis.z = z;
                        d.push_back(v);
                    }
        a.d = d.data();
        a.n = (int)d.size();
    }
};

std::vector<int> copy_of(const int *p, int n) { return std::vector<int>(p, p + n); }

}  // namespace

#define DRV __attribute__((visibility("default")))

extern "C" {

// S: the gft seed array (S[0] = count) or NULL; label in/out.
DRV int drv_oift_multi(const int *I, int xs, int ys, int zs, float adj_r, float per, const float *per_class,
                       int max_label, int *S, int *label)
{
    std::vector<int> img = copy_of(I, xs * ys * zs);
    Scene scn(img.data(), xs, ys, zs), lab(label, xs, ys, zs);
    Adj A(adj_r);
    if (per_class) gft::ift::OIFT_Multi_PerClass(&A.a, &scn.s, per_class, max_label, S, &lab.s);
    else gft::ift::OIFT_Multi(&A.a, &scn.s, per, S, &lab.s);
    return 0;
}

DRV int drv_relax_multi(const int *I, int xs, int ys, int zs, const float *spacing, float adj_r, float per,
                        const float *per_class, int max_label, int *S, int *label, int ntimes)
{
    std::vector<int> img = copy_of(I, xs * ys * zs);
    Scene scn(img.data(), xs, ys, zs, spacing), lab(label, xs, ys, zs, spacing);
    Adj A(adj_r);
    if (per_class) gft::ift::ORelax_1_Multi_PerClass(&A.a, &scn.s, per_class, max_label, S, &lab.s, ntimes);
    else gft::ift::ORelax_1_Multi(&A.a, &scn.s, per, S, &lab.s, ntimes);
    return 0;
}

// form 0: OIFT(A, scn, ...); 1: OIFT(W, A, scn, ...); 2: OIFT(Wx, Wy, Wz, scn, ...)
DRV int drv_oift(int form, const int *I, int *W, int *Wx, int *Wy, int *Wz, int xs, int ys, int zs, float adj_r,
                 float per, int *S, int *label)
{
    std::vector<int> img = copy_of(I, xs * ys * zs);
    Scene scn(img.data(), xs, ys, zs), lab(label, xs, ys, zs);
    Adj A(adj_r);
    if (form == 0) {
        gft::ift::OIFT(&A.a, &scn.s, per, S, &lab.s);
    } else if (form == 1) {
        Scene w(W, xs, ys, zs);
        gft::ift::OIFT(&w.s, &A.a, &scn.s, per, S, &lab.s);
    } else {
        Scene wx(Wx, xs, ys, zs), wy(Wy, xs, ys, zs), wz(Wz, xs, ys, zs);
        gft::ift::OIFT(&wx.s, &wy.s, &wz.s, &scn.s, per, S, &lab.s);
    }
    return 0;
}

// ORelax_1: W == Wx == NULL -> image form, W -> per-voxel form, Wx/Wy/Wz -> per-axis form.
DRV int drv_orelax_1(const int *I, int *W, int *Wx, int *Wy, int *Wz, int xs, int ys, int zs, const float *spacing,
                     float adj_r, float per, int *S, int *label, int ntimes)
{
    std::vector<int> img = copy_of(I, xs * ys * zs);
    Scene scn(img.data(), xs, ys, zs, spacing), lab(label, xs, ys, zs, spacing);
    Adj A(adj_r);
    if (Wx) {
        Scene wx(Wx, xs, ys, zs), wy(Wy, xs, ys, zs), wz(Wz, xs, ys, zs);
        gft::ift::ORelax_1(&wx.s, &wy.s, &wz.s, &scn.s, per, S, &lab.s, ntimes);
    } else if (W) {
        Scene w(W, xs, ys, zs);
        gft::ift::ORelax_1(&w.s, &A.a, &scn.s, per, S, &lab.s, ntimes);
    } else {
        gft::ift::ORelax_1(&A.a, &scn.s, per, S, &lab.s, ntimes);
    }
    return 0;
}

// gft::Scene32::GaussianBlur `passes` times; the result scenes come from Scene32::Create above.
DRV int drv_gaussian_blur(const int *I, int xs, int ys, int zs, const float *spacing, int passes, int *out)
{
    std::vector<int> img = copy_of(I, xs * ys * zs);
    Scene scn(img.data(), xs, ys, zs, spacing);
    gft::sScene32 *cur = &scn.s;
    for (int k = 0; k < passes; k++) {
        gft::sScene32 *b = gft::Scene32::GaussianBlur(cur);
        if (b->dx != cur->dx || b->dy != cur->dy || b->dz != cur->dz) return -1;   // spacing carried over
        if (cur != &scn.s) {
            std::free(cur->data);
            std::free(cur);
        }
        cur = b;
    }
    std::memcpy(out, cur->data, sizeof(int) * (size_t)(xs * ys * zs));
    if (cur != &scn.s) {
        std::free(cur->data);
        std::free(cur);
    }
    return 0;
}

}  // extern "C"
