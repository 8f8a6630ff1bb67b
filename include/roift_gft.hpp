// roift_gft.hpp -- link-level drop-in for the gft entry points on the hot path.
//
// libroift_gft.so defines, with the same qualified names, signatures and types
// as the reference's gft library, the functions its callers use on this path:
//   gft::ift::OIFT (A, scn, ...), (W, A, scn, ...) and (Wx, Wy, Wz, scn, ...)
//                                     (gft_ift.h:101-106, 125-139)
//   gft::ift::OIFT_Multi              (gft_ift.h:108-112)
//   gft::ift::OIFT_Multi_PerClass     (gft_ift.h:117-122)
//   gft::ift::ORelax_1_Multi          (gft_ift.h:451-456)
//   gft::ift::ORelax_1_Multi_PerClass (gft_ift.h:459-465)
//   gft::ift::ORelax_1 and its W / Wx,Wy,Wz forms (gft_ift.h:443-482)
//   gft::Scene32::GaussianBlur        (gft_filtering3.h:54), the CLI's preprocessing
// Each copies the scene to the GPU, runs libroift.so (include/roift.h) and writes
// the label scene back in place, exactly as the originals modify `label`.
//
// The record layouts below must match gft's (gft_scene32.h:20-27,
// gft_adjrel3.h:10-23) because callers allocate these records with gft's own
// constructors; only the fields this path reads are used (data, sizes, d, n).
// A program that includes the reference's gft headers does not include this
// file: it links -lroift_gft ahead of -lgft, so that these definitions are the ones the
// linker binds.
#ifndef ROIFT_GFT_HPP
#define ROIFT_GFT_HPP

namespace gft {

namespace AdjRel3 {
union sDisplacement3 {
    int data[4];
    struct {
        int x, y, z;
    } axis;
} __attribute__((aligned(16)));

struct sAdjRel3 {
    sDisplacement3 *d;
    int n;
};
}  // namespace AdjRel3

namespace Scene32 {
struct sScene32 {
    int *data;
    int ***array;
    int xsize, ysize, zsize;
    float dx, dy, dz;
    int maxval, minval, n;
    void *nii_hdr;
};
// gft's constructor (gft_scene32.h), defined by libgft, not by libroift_gft.so
sScene32 *Create(int xsize, int ysize, int zsize);
sScene32 *GaussianBlur(sScene32 *scn);
}  // namespace Scene32

typedef AdjRel3::sAdjRel3 sAdjRel3;
typedef Scene32::sScene32 sScene32;

namespace ift {
void OIFT(sAdjRel3 *A, sScene32 *scn, float per, int *S, sScene32 *label);
void OIFT(sScene32 *W, sAdjRel3 *A, sScene32 *scn, float per, int *S, sScene32 *label);
void OIFT(sScene32 *Wx, sScene32 *Wy, sScene32 *Wz, sScene32 *scn, float per, int *S, sScene32 *label);
void OIFT_Multi(sAdjRel3 *A, sScene32 *scn, float per, int *S, sScene32 *label);
void OIFT_Multi_PerClass(sAdjRel3 *A, sScene32 *scn, const float *per_class, int max_label, int *S,
                         sScene32 *label);
void ORelax_1_Multi(sAdjRel3 *A, sScene32 *scn, float per, int *S, sScene32 *label, int ntimes);
void ORelax_1_Multi_PerClass(sAdjRel3 *A, sScene32 *scn, const float *per_class, int max_label, int *S,
                             sScene32 *label, int ntimes);
void ORelax_1(sAdjRel3 *A, sScene32 *scn, float per, int *S, sScene32 *label, int ntimes);
void ORelax_1(sScene32 *W, sAdjRel3 *A, sScene32 *scn, float per, int *S, sScene32 *label, int ntimes);
void ORelax_1(sScene32 *Wx, sScene32 *Wy, sScene32 *Wz, sScene32 *scn, float per, int *S, sScene32 *label,
              int ntimes);
}  // namespace ift

}  // namespace gft

#endif
