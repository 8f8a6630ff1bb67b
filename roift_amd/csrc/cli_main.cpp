// cli_main.cpp -- roift_cli_main (in libroift.so) and roift_oiftrelax (its executable):
// the reference CLI's argv contract (oiftrelax.cpp:227-432) with every stage on the GPU:
//
//   roift_oiftrelax <volume> <file_seeds> <pol> <niter> <percentile> <output_file>
//                   [boundary_stride] [pol_file]
//
// read volume (negative minimum shifted to 0) -> seeds file -> boundary background seeds
// (stride, default 8, 0 = off) -> Gaussian blur x2 -> OIFT_Multi[_PerClass] -> ORelax_1_Multi
// [_PerClass] (niter) -> dilation_conditional(radius 1, percentile) when the largest label
// is <= 1 -> write the label volume (.scn, .nii, .nii.gz).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/roift.h"

namespace {

struct CliFail {};   // a stage failed; its message is already on stderr

struct Cli {
    roift_ctx *ctx = nullptr;
    std::vector<void *> dev;
    int32_t *h_label = nullptr;
    ~Cli()
    {
        for (void *p : dev) (void)hipFree(p);
        if (h_label) (void)hipHostFree(h_label);
        if (ctx) roift_destroy(ctx);
    }
    [[noreturn]] void die(const char *what, int rc, const char *msg)
    {
        std::fprintf(stderr, "roift_oiftrelax: %s failed (%d): %s\n", what, rc, msg);
        throw CliFail{};
    }
    void check(const char *what, int rc)
    {
        if (rc != ROIFT_OK) die(what, rc, roift_last_error(ctx));
    }
    void check_io(const char *what, int rc)
    {
        if (rc != ROIFT_OK) die(what, rc, roift_io_last_error());
    }
    void hip_check(const char *what, hipError_t e)
    {
        if (e != hipSuccess) die(what, (int)e, hipGetErrorString(e));
    }
    template <typename T>
    T *dmalloc(size_t bytes)
    {
        void *p = nullptr;
        hip_check("hipMalloc", hipMalloc(&p, bytes));
        dev.push_back(p);
        return (T *)p;
    }
};

struct Timer {
    std::vector<std::pair<std::string, double>> ev;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    void mark(const std::string &name)
    {
        const auto t1 = std::chrono::steady_clock::now();
        ev.emplace_back(name, std::chrono::duration<double, std::milli>(t1 - t0).count());
        t0 = t1;
    }
};

void usage()
{
    std::printf("usage:\n");
    std::printf("roift_oiftrelax <volume> <file_seeds> <pol> <niter> <percentile> <output_file> [boundary_stride] "
                "[pol_file]\n");
    std::printf("\t pol.............. Global boundary polarity [-1.0, 1.0] (ignored if pol_file given)\n");
    std::printf("\t niter............ Relaxation iterations\n");
    std::printf("\t percentile...... Dilation percentile (binary mode only)\n");
    std::printf("\t output_file..... Output label file (.scn, .nii, .nii.gz)\n");
    std::printf("\t boundary_stride. Stride for auto boundary bg seeds (default=8, 0=off)\n");
    std::printf("\t pol_file........ Per-class polarity file (optional): <n_labels> then pol_0 ... pol_{n-1}\n");
}

int run(Cli &cl, int argc, char **argv)
{
    auto check = [&](const char *w, int rc) { cl.check(w, rc); };
    auto check_io = [&](const char *w, int rc) { cl.check_io(w, rc); };
    auto hip_check = [&](const char *w, hipError_t e) { cl.hip_check(w, e); };
    Timer tm;
    check("roift_create", roift_create(&cl.ctx, 0));
    roift_ctx *g_ctx = cl.ctx;

    // volume (the CLI shifts a negative minimum to 0: oiftrelax.cpp:263-268)
    roift_volume_info info{};
    check_io("reading the volume header", roift_volume_info_read(argv[1], &info));
    const long long n = (long long)info.xsize * info.ysize * info.zsize;
    void *d_img = cl.dmalloc<void>((size_t)n * 4);
    int dtype = ROIFT_I16, shift = 0;
    if (roift_volume_read_dev(g_ctx, argv[1], d_img, ROIFT_I16, ROIFT_READ_SHIFT_MIN, &shift, &info) != ROIFT_OK) {
        dtype = ROIFT_I32;   // values beyond int16
        check("reading the volume", roift_volume_read_dev(g_ctx, argv[1], d_img, ROIFT_I32, ROIFT_READ_SHIFT_MIN,
                                                          &shift, &info));
    }
    tm.mark("read volume");

    const float pol = (float)std::atof(argv[3]);
    const int niter = std::atoi(argv[4]);
    const int percentile = std::atoi(argv[5]);
    const char *output_file = argv[6];
    const int boundary_stride = argc >= 8 ? std::atoi(argv[7]) : 8;
    std::vector<float> per_class(1024);
    int n_classes = 0;
    if (argc >= 9) {
        int rc = roift_polarity_read(argv[8], per_class.data(), (int)per_class.size(), &n_classes);
        if (rc != ROIFT_OK) {
            std::fprintf(stderr, "Warning: polarity file not used: %s\n", roift_io_last_error());
            n_classes = 0;
        } else {
            std::printf("Per-class polarity loaded: %d classes\n", n_classes);
        }
    }

    // seeds file -> seed list and label map (file order; a later line wins a voxel)
    int count = 0, clamped = 0;
    check_io("reading the seeds",
             roift_seeds_read(argv[2], info.xsize, info.ysize, info.zsize, nullptr, nullptr, 0, &count, &clamped));
    std::vector<int32_t> seeds((size_t)std::max(count, 1)), slab((size_t)std::max(count, 1));
    check_io("reading the seeds", roift_seeds_read(argv[2], info.xsize, info.ysize, info.zsize, seeds.data(),
                                                   slab.data(), count, &count, &clamped));
    if (clamped) std::printf("Warning: negative seed labels were clamped to 0 (background).\n");
    hip_check("hipHostMalloc", hipHostMalloc((void **)&cl.h_label, (size_t)n * 4, hipHostMallocDefault));
    int32_t *h_label = cl.h_label;
    std::fill(h_label, h_label + n, -1);
    for (int i = 0; i < count; i++) h_label[seeds[i]] = slab[i];
    const int face_cap = 2 * (info.xsize * info.ysize + info.ysize * info.zsize + info.xsize * info.zsize) + 8;
    int32_t *d_label = cl.dmalloc<int32_t>((size_t)n * 4);
    int32_t *d_seeds = cl.dmalloc<int32_t>((size_t)(count + face_cap) * 4);
    hip_check("copy labels", hipMemcpy(d_label, h_label, (size_t)n * 4, hipMemcpyHostToDevice));
    if (count > 0) hip_check("copy seeds", hipMemcpy(d_seeds, seeds.data(), (size_t)count * 4, hipMemcpyHostToDevice));
    // the scene's spacing reaches the blur taps and the relaxation's Dpq, as the reference's
    // fscn carries it (gft_filtering3.cpp:390-392)
    roift_problem pb{info.xsize, info.ysize, info.zsize, 1.0f, 0, 0.0f, nullptr, -1, info.dx, info.dy, info.dz};
    int total = count;
    if (boundary_stride > 0) {   // appended after the file's seeds, in raster order (oiftrelax.cpp:357-364)
        int added = 0;
        check("boundary seeds", roift_boundary_seeds_dev(g_ctx, &pb, d_label, boundary_stride, 0, d_seeds, count,
                                                         d_seeds + count, face_cap, &added));
        total = count + added;
        std::printf("Boundary seeds: %d bg seeds added (stride=%d), total seeds: %d\n", added, boundary_stride, total);
    }
    tm.mark("seeds");

    check("GaussianBlur", roift_gaussian_blur_dev(g_ctx, &pb, d_img, dtype, d_img, 2));
    tm.mark("GaussianBlur x2");

    if (n_classes > 0) {
        pb.use_per_class = 1;
        pb.h_per_class = per_class.data();
        pb.max_label = n_classes - 1;
    } else {
        pb.per = (float)(pol * 100.0);   // OIFT_Multi(A, fscn, pol * 100.0, S, label)
    }
    check("OIFT + relaxation", roift_pipeline_dev(g_ctx, &pb, d_img, dtype, d_seeds, total, d_label, niter, nullptr,
                                                  nullptr));
    tm.mark("OIFT + ORelax_1_Multi");

    hip_check("copy labels", hipMemcpy(h_label, d_label, (size_t)n * 4, hipMemcpyDeviceToHost));
    int max_label = INT32_MIN;
    for (long long p = 0; p < n; p++) max_label = std::max(max_label, h_label[p]);
    if (max_label <= 1) {
        roift_problem dp = pb;
        dp.adj_radius = 1.0f;
        dp.use_per_class = 0;
        int set = 0;
        check("dilation_conditional",
              roift_dilation_conditional_dev(g_ctx, &dp, d_img, dtype, d_label, (float)percentile, &set));
        hip_check("copy labels", hipMemcpy(h_label, d_label, (size_t)n * 4, hipMemcpyDeviceToHost));
        tm.mark("Dilation Conditional");
    } else {
        std::printf("Skipping binary dilation post-process for multi-label result (max label=%d).\n", max_label);
    }

    roift_volume_info out = info;
    check_io("writing the labels", roift_volume_write(output_file, h_label, &out, argv[1]));
    tm.mark("write labels");

    std::printf("Stage timings (ms):\n");
    for (auto &e : tm.ev) std::printf("  %-28s %10.3f\n", e.first.c_str(), e.second);
    std::fflush(stdout);
    return 0;
}

}  // namespace

// 0 on success (and for the usage text), 1 on any failure (the message goes to stderr).
extern "C" __attribute__((visibility("default"))) int roift_cli_main(int argc, char **argv)
{
    if (argc < 7) {
        usage();
        return 0;
    }
    Cli cl;
    try {
        return run(cl, argc, argv);
    } catch (const CliFail &) {
        return 1;
    }
}
