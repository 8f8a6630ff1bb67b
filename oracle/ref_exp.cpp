// ref_exp.cpp -- extern "C" wrappers (written for this project) around the
// reference's experiment variants of the hot path, compiled from the experiment
// sources where they lie (oracle/Makefile passes their paths as EXP_GRAD and
// EXP_RBF).  Each experiment is a self-contained program whose functions are
// file-static, so each is included into a namespace of its own here.
//
// TEST INFRASTRUCTURE ONLY (linked into oracle/_ref/libroift_ref.so): lets the
// tests compare the oracle's restatements of
//   OIFT_Multi_GradWeight / ORelax_1_Multi_GradWeight (exp_gradient_weight.cpp:196-504)
//   ORelax_1_Multi_GaussRBF, estimate_sigma          (exp_gaussian_rbf_relax.cpp:194-427)
//   compute_bfs_edt, inject_boundary_seeds_geodesic  (exp_geodesic_seeds.cpp:165-283)
//   dilation_conditional, inject_boundary_seeds      (oiftrelax.cpp:87-222)
// with the reference's own code (the CLI, oiftrelax.cpp, is included the same way; its
// main and the experiments' become namespaced functions that are never called).
#include "gft.h"
#include "gft_gradient3.h"
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <map>
#include <queue>
#include <string>
#include <vector>

namespace exp_grad {
#include EXP_GRAD
}
namespace exp_geo {
#include EXP_GEO
}
namespace cli {
#include CLI_MAIN
}
namespace exp_rbf {
#include EXP_RBF
}

// the voxel spacing ref_set_spacing gave (ref_capi.cpp)
extern "C" void ref_get_spacing(float *out3);

namespace {

gft::sScene32 *scene_of(const int *data, int xs, int ys, int zs)
{
    gft::sScene32 *s = gft::Scene32::Create(xs, ys, zs);
    std::memcpy(s->data, data, sizeof(int) * (size_t)s->n);
    float sp[3];
    ref_get_spacing(sp);
    s->dx = sp[0];
    s->dy = sp[1];
    s->dz = sp[2];
    return s;
}

int *S_of(const int *S)
{
    int *c = (int *)std::calloc((size_t)S[0] + 1, sizeof(int));
    std::memcpy(c, S, sizeof(int) * ((size_t)S[0] + 1));
    return c;
}

}  // namespace

#define REF_API __attribute__((visibility("default")))

extern "C" {

// OIFT_Multi_GradWeight; label in/out.
REF_API int ref_oift_gradweight(const int *I, const int *G, int xs, int ys, int zs, float adj_r, float per,
                                float alpha, const int *S, int *label)
{
    gft::sAdjRel3 *A = gft::AdjRel3::Spheric(adj_r);
    gft::sScene32 *scn = scene_of(I, xs, ys, zs), *g = scene_of(G, xs, ys, zs), *lab = scene_of(label, xs, ys, zs);
    int *Sc = S_of(S);
    exp_grad::OIFT_Multi_GradWeight(A, scn, g, per, alpha, Sc, lab);
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    gft::Scene32::Destroy(&g);
    gft::Scene32::Destroy(&scn);
    gft::AdjRel3::Destroy(&A);
    return 0;
}

// ORelax_1_Multi_GradWeight; label in = OIFT labels, out = relaxed labels.
REF_API int ref_relax_gradweight(const int *I, const int *G, int xs, int ys, int zs, float adj_r, float per,
                                 float alpha, const int *S, int *label, int ntimes)
{
    gft::sAdjRel3 *A = gft::AdjRel3::Spheric(adj_r);
    gft::sScene32 *scn = scene_of(I, xs, ys, zs), *g = scene_of(G, xs, ys, zs), *lab = scene_of(label, xs, ys, zs);
    int *Sc = S_of(S);
    exp_grad::ORelax_1_Multi_GradWeight(A, scn, g, per, alpha, Sc, lab, ntimes);
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    gft::Scene32::Destroy(&g);
    gft::Scene32::Destroy(&scn);
    gft::AdjRel3::Destroy(&A);
    return 0;
}

// ORelax_1_Multi_GaussRBF; label in/out.
REF_API int ref_relax_rbf(const int *I, int xs, int ys, int zs, float adj_r, float per, float sigma, const int *S,
                          int *label, int ntimes)
{
    gft::sAdjRel3 *A = gft::AdjRel3::Spheric(adj_r);
    gft::sScene32 *scn = scene_of(I, xs, ys, zs), *lab = scene_of(label, xs, ys, zs);
    int *Sc = S_of(S);
    exp_rbf::ORelax_1_Multi_GaussRBF(A, scn, per, sigma, Sc, lab, ntimes);
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    gft::Scene32::Destroy(&scn);
    gft::AdjRel3::Destroy(&A);
    return 0;
}

REF_API float ref_estimate_sigma(const int *I, int xs, int ys, int zs)
{
    gft::sScene32 *scn = scene_of(I, xs, ys, zs);
    const float s = exp_rbf::estimate_sigma(scn);
    gft::Scene32::Destroy(&scn);
    return s;
}

// compute_bfs_edt; dist receives N ints (INT_MAX = unreached).  S is gft's seed array.
REF_API int ref_bfs_edt(const int *label, int xs, int ys, int zs, const int *S, int *dist)
{
    gft::sScene32 *lab = scene_of(label, xs, ys, zs);
    const std::vector<int> d = exp_geo::compute_bfs_edt(lab, S);
    std::memcpy(dist, d.data(), sizeof(int) * d.size());
    gft::Scene32::Destroy(&lab);
    return 0;
}

// inject_boundary_seeds_geodesic (edt computed by compute_bfs_edt from S, as the experiment
// does); label in/out; S_out (capacity cap) receives gft's seed array; returns the count added.
REF_API int ref_boundary_seeds_geodesic(int *label, int xs, int ys, int zs, const int *S, int stride, int *S_out,
                                        int cap)
{
    gft::sScene32 *lab = scene_of(label, xs, ys, zs);
    int *Sc = S_of(S);
    const std::vector<int> d = exp_geo::compute_bfs_edt(lab, Sc);
    const int added = exp_geo::inject_boundary_seeds_geodesic(lab, Sc, Sc[0], stride, d);
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    if (Sc[0] + 1 <= cap) std::memcpy(S_out, Sc, sizeof(int) * ((size_t)Sc[0] + 1));
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    return added;
}

// the CLI's inject_boundary_seeds; label in/out; S_out as above.
REF_API int ref_boundary_seeds(int *label, int xs, int ys, int zs, const int *S, int stride, int *S_out, int cap)
{
    gft::sScene32 *lab = scene_of(label, xs, ys, zs);
    int *Sc = S_of(S);
    const int added = cli::inject_boundary_seeds(lab, Sc, Sc[0], stride);
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    if (Sc[0] + 1 <= cap) std::memcpy(S_out, Sc, sizeof(int) * ((size_t)Sc[0] + 1));
    std::free(Sc);
    gft::Scene32::Destroy(&lab);
    return added;
}

// the CLI's dilation_conditional(fscn, label, radius, percentile); label in/out.
REF_API int ref_dilation_conditional(const int *I, int *label, int xs, int ys, int zs, float radius, float percentile)
{
    gft::sScene32 *scn = scene_of(I, xs, ys, zs), *lab = scene_of(label, xs, ys, zs);
    cli::dilation_conditional(scn, lab, radius, percentile);
    std::memcpy(label, lab->data, sizeof(int) * (size_t)lab->n);
    gft::Scene32::Destroy(&lab);
    gft::Scene32::Destroy(&scn);
    return 0;
}

// the reference CLI itself (oiftrelax.cpp:227-432), run in-process on files
REF_API int ref_cli_main(int argc, char **argv) { return cli::main(argc, argv); }

}  // extern "C"
