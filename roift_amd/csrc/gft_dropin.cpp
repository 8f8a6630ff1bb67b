// gft_dropin.cpp -- gft-signature entry points over the C-ABI (libroift_gft.so).
//
// Same behaviour contract as the reference functions they replace
// (gft_ift.cpp:828-1271, 8498-9498; GaussianBlur, gft_filtering3.cpp:385-397): `label` is modified in place; a NULL S
// seeds every labelled voxel in raster order (gft_ift.cpp:944-948); where the
// reference would call gft::Error (exit(-1)), this prints the library's error
// and exits with the same status, since these void signatures cannot report it.
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "../../include/roift.h"
#include "../../include/roift_gft.hpp"

namespace {

roift_ctx *context()
{
    static roift_ctx *ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        if (roift_create(&ctx, 0) != ROIFT_OK) {
            std::fprintf(stderr, "roift: no HIP device for the gft drop-in\n");
            std::exit(-1);
        }
    });
    return ctx;
}

float radius_of(const gft::sAdjRel3 *A)
{
    // Spheric(1.0) has 7 displacements, Spheric(1.75) has 27 (gft_adjrel3.cpp:47-85)
    if (A->n == 7) return 1.0f;
    if (A->n == 27) return 1.75f;
    std::fprintf(stderr, "roift: adjacency with %d displacements is not supported\n", A->n);
    std::exit(-1);
}

std::vector<int> seeds_of(const int *S, const gft::sScene32 *label)
{
    std::vector<int> s;
    if (S) {
        s.assign(S + 1, S + 1 + S[0]);
    } else {
        for (int p = 0; p < label->n; p++)
            if (label->data[p] != -1) s.push_back(p);
    }
    return s;
}

void check(int rc)
{
    if (rc != ROIFT_OK) {
        std::fprintf(stderr, "roift error %d: %s\n", rc, roift_last_error(context()));
        std::exit(-1);
    }
}

roift_problem problem(const gft::sAdjRel3 *A, const gft::sScene32 *scn, int use_pc, float per, const float *pc,
                      int max_label)
{
    return roift_problem{scn->xsize, scn->ysize, scn->zsize, radius_of(A), use_pc, per, pc, max_label,
                         scn->dx, scn->dy, scn->dz};
}

}  // namespace

namespace gft {
namespace ift {

__attribute__((visibility("default"))) void OIFT_Multi(sAdjRel3 *A, sScene32 *scn, float per, int *S, sScene32 *label)
{
    const std::vector<int> s = seeds_of(S, label);
    const roift_problem pb = problem(A, scn, 0, per, nullptr, -1);
    check(roift_oift_fmax(context(), &pb, scn->data, s.data(), (int)s.size(), label->data, nullptr));
}

__attribute__((visibility("default"))) void OIFT_Multi_PerClass(sAdjRel3 *A, sScene32 *scn, const float *per_class,
                                                                int max_label, int *S, sScene32 *label)
{
    const std::vector<int> s = seeds_of(S, label);
    const roift_problem pb = problem(A, scn, 1, 0.0f, per_class, max_label);
    check(roift_oift_fmax(context(), &pb, scn->data, s.data(), (int)s.size(), label->data, nullptr));
}

// OIFT(A, scn, per, S, label) (gft_ift.cpp:828-908): the two-label outer cut polarises by
// label > 0 ? per : -per, which for the labels >= 0 a popped voxel carries is OIFT_Multi's rule.
__attribute__((visibility("default"))) void OIFT(sAdjRel3 *A, sScene32 *scn, float per, int *S, sScene32 *label)
{
    const std::vector<int> s = seeds_of(S, label);
    const roift_problem pb = problem(A, scn, 0, per, nullptr, -1);
    check(roift_oift_fmax(context(), &pb, scn->data, s.data(), (int)s.size(), label->data, nullptr));
}

// OIFT(W, A, scn, per, S, label) (gft_ift.cpp:1086-1167): base W_p + W_q.
__attribute__((visibility("default"))) void OIFT(sScene32 *W, sAdjRel3 *A, sScene32 *scn, float per, int *S,
                                                 sScene32 *label)
{
    const std::vector<int> s = seeds_of(S, label);
    const roift_problem pb = problem(A, scn, 0, per, nullptr, -1);
    check(roift_oift_w(context(), &pb, scn->data, W->data, nullptr, nullptr, nullptr, s.data(), (int)s.size(),
                       label->data, nullptr));
}

// OIFT(Wx, Wy, Wz, scn, per, S, label) (gft_ift.cpp:1170-1271): per-axis bases on Spheric(1.0).
__attribute__((visibility("default"))) void OIFT(sScene32 *Wx, sScene32 *Wy, sScene32 *Wz, sScene32 *scn, float per,
                                                 int *S, sScene32 *label)
{
    static gft::AdjRel3::sDisplacement3 d6[7];
    static sAdjRel3 six{d6, 7};   // gft_ift.cpp:1191
    const std::vector<int> s = seeds_of(S, label);
    const roift_problem pb = problem(&six, scn, 0, per, nullptr, -1);
    check(roift_oift_w(context(), &pb, scn->data, nullptr, Wx->data, Wy->data, Wz->data, s.data(), (int)s.size(),
                       label->data, nullptr));
}

__attribute__((visibility("default"))) void ORelax_1_Multi(sAdjRel3 *A, sScene32 *scn, float per, int *S,
                                                           sScene32 *label, int ntimes)
{
    if (S == nullptr || S[0] <= 0) {   // the binary fallback (gft_ift.cpp:8678-8682)
        ORelax_1(A, scn, per, S, label, ntimes);
        return;
    }
    const std::vector<int> s = seeds_of(S, label);
    const roift_problem pb = problem(A, scn, 0, per, nullptr, -1);
    check(roift_relax(context(), &pb, scn->data, s.data(), (int)s.size(), label->data, ntimes, nullptr));
}

__attribute__((visibility("default"))) void ORelax_1_Multi_PerClass(sAdjRel3 *A, sScene32 *scn, const float *per_class,
                                                                    int max_label, int *S, sScene32 *label, int ntimes)
{
    if (S == nullptr || S[0] <= 0) return;  // gft_ift.cpp:8942-8943
    const std::vector<int> s = seeds_of(S, label);
    const roift_problem pb = problem(A, scn, 1, 0.0f, per_class, max_label);
    check(roift_relax(context(), &pb, scn->data, s.data(), (int)s.size(), label->data, ntimes, nullptr));
}

namespace {

// S == NULL reads as no seeds here (the reference dereferences S[0]).
void relax_binary(sAdjRel3 *A, sScene32 *scn, const int *W, const int *Wx, const int *Wy, const int *Wz, float per,
                  int *S, sScene32 *label, int ntimes)
{
    const std::vector<int> s = S ? std::vector<int>(S + 1, S + 1 + S[0]) : std::vector<int>();
    const roift_problem pb = problem(A, scn, 0, per, nullptr, -1);
    check(roift_relax_binary(context(), &pb, scn->data, W, Wx, Wy, Wz, s.data(), (int)s.size(), label->data, ntimes,
                             nullptr));
}

}  // namespace

__attribute__((visibility("default"))) void ORelax_1(sAdjRel3 *A, sScene32 *scn, float per, int *S, sScene32 *label,
                                                     int ntimes)
{
    relax_binary(A, scn, nullptr, nullptr, nullptr, nullptr, per, S, label, ntimes);
}

__attribute__((visibility("default"))) void ORelax_1(sScene32 *W, sAdjRel3 *A, sScene32 *scn, float per, int *S,
                                                     sScene32 *label, int ntimes)
{
    relax_binary(A, scn, W->data, nullptr, nullptr, nullptr, per, S, label, ntimes);
}

__attribute__((visibility("default"))) void ORelax_1(sScene32 *Wx, sScene32 *Wy, sScene32 *Wz, sScene32 *scn,
                                                     float per, int *S, sScene32 *label, int ntimes)
{
    static gft::AdjRel3::sDisplacement3 d6[7];
    static sAdjRel3 six{d6, 7};   // the per-axis form always uses Spheric(1.0) (gft_ift.cpp:9355)
    relax_binary(&six, scn, nullptr, Wx->data, Wy->data, Wz->data, per, S, label, ntimes);
}

}  // namespace ift

namespace Scene32 {

// The blurred scene is allocated by gft's own constructor (resolved from libgft at the
// caller's link), so gft::Scene32::Destroy frees it as usual.
__attribute__((visibility("default"))) sScene32 *GaussianBlur(sScene32 *scn)
{
    sScene32 *blur = Create(scn->xsize, scn->ysize, scn->zsize);
    blur->dx = scn->dx;
    blur->dy = scn->dy;
    blur->dz = scn->dz;
    const roift_problem pb{scn->xsize, scn->ysize, scn->zsize, 1.0f, 0, 0.0f, nullptr, -1, scn->dx, scn->dy, scn->dz};
    check(roift_gaussian_blur(context(), &pb, scn->data, blur->data, 1));
    return blur;
}

}  // namespace Scene32
}  // namespace gft
