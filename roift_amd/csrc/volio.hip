// volio.hip -- volume, seed and polarity files for the CLI's argv contract
// (oiftrelax.cpp:258-351), and the volume's way onto the device.
//
// Formats, written from their published layouts (not from gft's code):
//   SCN      gft's raw scene: "SCN\n<xs> <ys> <zs>\n<dx> <dy> <dz>\n<bits>" and one end-of-line
//            character, then xs*ys*zs little-endian samples of 8 (unsigned), 16 (unsigned)
//            or 32 (signed) bits (read: gft_scene32.cpp:334-419; write, bits chosen from
//            the maximum: :602-664)
//   NIfTI-1  the 348-byte header + data (.nii, .nii.gz, .hdr/.img), types uint8, int8,
//            int16, uint16, int32, float32 -- the ones gft::Scene32::ReadNifti1 accepts
//            (gft_scene32.cpp:525-597); float32 is scaled by 10000 / max and rounded with
//            gft's ROUND, as gft does.  Labels are written as int32 NIfTI-1, with the
//            input's header fields kept when the input was NIfTI (gft clones the header).
//   seeds    "<n>" then n lines "<x> <y> <z> <id> <label>": invalid voxels are skipped,
//            negative labels become 0 (oiftrelax.cpp:323-351)
//   polarity "<n_labels>" then n values in [-1, 1], scaled by 100.0f (oiftrelax.cpp:276-310)
//
// The device read stages the file's raw samples through two pinned host buffers
// (double-buffered: the next chunk is read from disk while the previous one is copied),
// and converts them to the requested int16/int32 layout in a kernel.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "ctx.hpp"

namespace roift {
namespace {

thread_local std::string g_io_error;

struct IoError {
    int code;
    std::string msg;
};

[[noreturn]] void io_fail(int code, const std::string &msg) { throw IoError{code, msg}; }

bool ends_with(const std::string &s, const char *suf)
{
    const size_t n = std::strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; i++)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != std::tolower((unsigned char)suf[i])) return false;
    return true;
}

// A gzip-or-plain input file (zlib reads uncompressed files transparently).
struct InFile {
    gzFile f = nullptr;
    explicit InFile(const std::string &path)
    {
        f = gzopen(path.c_str(), "rb");
        if (!f) io_fail(ROIFT_EINVAL, "cannot open " + path);
        gzbuffer(f, 1 << 20);
    }
    ~InFile()
    {
        if (f) gzclose(f);
    }
    void read(void *dst, size_t bytes, const char *what)
    {
        char *p = (char *)dst;
        while (bytes > 0) {
            const unsigned chunk = (unsigned)std::min<size_t>(bytes, 1u << 30);
            const int got = gzread(f, p, chunk);
            if (got <= 0) io_fail(ROIFT_EINVAL, std::string("truncated file reading ") + what);
            p += got;
            bytes -= (size_t)got;
        }
    }
    void skip_to(long long off)
    {
        if (gzseek(f, (z_off_t)off, SEEK_SET) < 0) io_fail(ROIFT_EINVAL, "cannot seek to the data");
    }
    // a whitespace-separated token, as fscanf("%s") / ("%d") read them
    std::string token()
    {
        std::string t;
        int ch;
        while ((ch = gzgetc(f)) != -1 && std::isspace(ch)) {}
        while (ch != -1 && !std::isspace(ch)) {
            t.push_back((char)ch);
            ch = gzgetc(f);
        }
        if (ch != -1) gzungetc(ch, f);
        return t;
    }
    int getc() { return gzgetc(f); }
};

// ---- NIfTI-1 header fields (offsets of the published 348-byte layout) ----------------
enum : int { NII_UINT8 = 2, NII_INT16 = 4, NII_INT32 = 8, NII_FLOAT32 = 16, NII_INT8 = 256, NII_UINT16 = 512 };

template <typename T>
T hget(const unsigned char *h, int off, bool swap)
{
    T v;
    std::memcpy(&v, h + off, sizeof(T));
    if (swap) {
        unsigned char b[sizeof(T)];
        std::memcpy(b, &v, sizeof(T));
        std::reverse(b, b + sizeof(T));
        std::memcpy(&v, b, sizeof(T));
    }
    return v;
}

template <typename T>
void hset(unsigned char *h, int off, T v)
{
    std::memcpy(h + off, &v, sizeof(T));
}

struct Header {
    int format = 0;         // 1 SCN, 2 NIfTI-1
    int xs = 0, ys = 0, zs = 0;
    float dx = 1, dy = 1, dz = 1;
    int stored = 0;         // ROIFT_U8 / I8 / U16 / I16 / I32 / F32
    int esize = 0;          // bytes per stored sample
    bool swap = false;      // NIfTI written with the other byte order
    std::string data_path;  // file holding the samples
    long long data_off = 0; // offset of the first sample in data_path
    unsigned char nii[348] = {};   // the NIfTI header as read (native byte order after parse)
    long long n() const { return (long long)xs * ys * zs; }
};

void read_header(const std::string &path, Header &h)
{
    if (ends_with(path, ".nii") || ends_with(path, ".nii.gz") || ends_with(path, ".hdr") || ends_with(path, ".img")) {
        h.format = 2;
        std::string hdr = path, img = path;
        const bool pair = ends_with(path, ".hdr") || ends_with(path, ".img");
        if (pair) {
            const std::string base = path.substr(0, path.size() - 4);
            hdr = base + ".hdr";
            img = base + ".img";
        }
        InFile f(hdr);
        unsigned char raw[348];
        f.read(raw, 348, "the NIfTI header");
        int sz = hget<int>(raw, 0, false);
        h.swap = false;
        if (sz != 348) {
            if (hget<int>(raw, 0, true) != 348) io_fail(ROIFT_EINVAL, path + ": not a NIfTI-1 header");
            h.swap = true;
        }
        if (std::memcmp(raw + 344, "n+1", 4) != 0 && std::memcmp(raw + 344, "ni1", 4) != 0)
            io_fail(ROIFT_EUNSUPPORTED, path + ": not NIfTI-1 (magic)");
        short dim[8];
        for (int i = 0; i < 8; i++) dim[i] = hget<short>(raw, 40 + 2 * i, h.swap);
        const short dt = hget<short>(raw, 70, h.swap);
        float pix[8];
        for (int i = 0; i < 8; i++) pix[i] = hget<float>(raw, 76 + 4 * i, h.swap);
        const float vox_offset = hget<float>(raw, 108, h.swap);
        if (dim[0] < 3 || dim[0] > 4) io_fail(ROIFT_EUNSUPPORTED, path + ": only 3-D (or 4-D, first volume) images");
        h.xs = std::max<int>(1, dim[1]);
        h.ys = std::max<int>(1, dim[2]);
        h.zs = std::max<int>(1, dim[3]);
        // nifti2_io sets a zero or non-finite grid spacing of a used axis to 1.0 (nifti2_io.c:3670-3673)
        for (int i = 1; i <= dim[0] && i < 8; i++)
            if (pix[i] == 0.0f || !std::isfinite(pix[i])) pix[i] = 1.0f;
        h.dx = pix[1];
        h.dy = pix[2];
        h.dz = pix[3];
        switch (dt) {
        case NII_UINT8: h.stored = ROIFT_U8; h.esize = 1; break;
        case NII_INT8: h.stored = ROIFT_I8; h.esize = 1; break;
        case NII_INT16: h.stored = ROIFT_I16; h.esize = 2; break;
        case NII_UINT16: h.stored = ROIFT_U16; h.esize = 2; break;
        case NII_INT32: h.stored = ROIFT_I32; h.esize = 4; break;
        case NII_FLOAT32: h.stored = ROIFT_F32; h.esize = 4; break;
        default: io_fail(ROIFT_EUNSUPPORTED, path + ": NIfTI datatype " + std::to_string(dt) + " is not supported");
        }
        h.data_path = pair ? img : path;
        h.data_off = pair ? (long long)vox_offset : std::max(352LL, (long long)vox_offset);
        // keep the header in native byte order for cloning into the output
        std::memcpy(h.nii, raw, 348);
        if (h.swap) {
            auto sw = [&](int off, int sz2) { std::reverse(h.nii + off, h.nii + off + sz2); };
            sw(0, 4); sw(32, 4); sw(36, 2);
            for (int i = 0; i < 8; i++) sw(40 + 2 * i, 2);
            for (int i = 0; i < 3; i++) sw(56 + 4 * i, 4);
            sw(68, 2); sw(70, 2); sw(72, 2); sw(74, 2);
            for (int i = 0; i < 8; i++) sw(76 + 4 * i, 4);
            for (int off : {108, 112, 116}) sw(off, 4);
            sw(120, 2);
            for (int off : {124, 128, 132, 136, 140, 144}) sw(off, 4);
            sw(252, 2); sw(254, 2);
            for (int i = 0; i < 18; i++) sw(256 + 4 * i, 4);
        }
    } else if (ends_with(path, ".scn")) {
        h.format = 1;
        InFile f(path);
        if (f.token() != "SCN") io_fail(ROIFT_EINVAL, path + ": not an SCN file");
        try {
            h.xs = std::stoi(f.token());
            h.ys = std::stoi(f.token());
            h.zs = std::stoi(f.token());
            h.dx = std::stof(f.token());
            h.dy = std::stof(f.token());
            h.dz = std::stof(f.token());
            const int bits = std::stoi(f.token());
            h.stored = bits == 8 ? ROIFT_U8 : (bits == 16 ? ROIFT_U16 : ROIFT_I32);
            h.esize = bits == 8 ? 1 : (bits == 16 ? 2 : 4);
        } catch (const std::exception &) {
            io_fail(ROIFT_EINVAL, path + ": malformed SCN header");
        }
        // the samples start one character after the bit count (its end of line)
        f.getc();
        h.data_off = (long long)gztell(f.f);
        h.data_path = path;
    } else {
        io_fail(ROIFT_EUNSUPPORTED, path + ": unsupported extension (.scn, .nii, .nii.gz, .hdr/.img)");
    }
    if (h.xs <= 0 || h.ys <= 0 || h.zs <= 0 || h.n() >= (1LL << 31))
        io_fail(ROIFT_EINVAL, path + ": bad volume size");
}

int32_t sample(const unsigned char *p, int stored, bool swap)
{
    unsigned char b[4];
    const int sz = stored == ROIFT_U8 || stored == ROIFT_I8 ? 1 : (stored == ROIFT_I16 || stored == ROIFT_U16 ? 2 : 4);
    std::memcpy(b, p, sz);
    if (swap) std::reverse(b, b + sz);
    switch (stored) {
    case ROIFT_U8: return b[0];
    case ROIFT_I8: return (int8_t)b[0];
    case ROIFT_U16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
    case ROIFT_I16: { int16_t v; std::memcpy(&v, b, 2); return v; }
    default: { int32_t v; std::memcpy(&v, b, 4); return v; }
    }
}

float sample_f32(const unsigned char *p, bool swap)
{
    unsigned char b[4];
    std::memcpy(b, p, 4);
    if (swap) std::reverse(b, b + 4);
    float v;
    std::memcpy(&v, b, 4);
    return v;
}

// gft's float conversion factor: 10000.0 / max(0, samples), stored as float
float f32_factor(float mx) { return (float)(10000.0 / (double)mx); }

int32_t gft_round_f(float x) { return x < 0 ? (int32_t)((double)x - 0.5) : (int32_t)((double)x + 0.5); }

template <class F>
int io_guard(F &&f)
{
    try {
        return f();
    } catch (const IoError &e) {
        g_io_error = e.msg;
        return e.code;
    } catch (const Error &e) {
        g_io_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_io_error = "out of host memory";
        return ROIFT_ENOMEM;
    }
}

// ---- device conversion ------------------------------------------------------------

template <typename FT, bool SWAP>
__device__ __forceinline__ FT load_sample(const FT *raw, long long i)
{
    FT v = raw[i];
    if constexpr (SWAP) {
        if constexpr (sizeof(FT) == 2) {
            const uint16_t u = __builtin_bit_cast(uint16_t, v);
            v = __builtin_bit_cast(FT, (uint16_t)((u >> 8) | (u << 8)));
        } else if constexpr (sizeof(FT) == 4) {
            v = __builtin_bit_cast(FT, __builtin_bswap32(__builtin_bit_cast(uint32_t, v)));
        }
    }
    return v;
}

// stored sample -> int32 as gft reads it (float32: ROUND(v * factor) in FP32)
template <typename FT, bool SWAP>
__device__ __forceinline__ int32_t to_i32(const FT *raw, long long i, float factor)
{
    const FT v = load_sample<FT, SWAP>(raw, i);
    if constexpr (std::is_floating_point<FT>::value) {
        const float x = (float)v * factor;
        return x < 0.0f ? (int32_t)((double)x - 0.5) : (int32_t)((double)x + 0.5);
    } else {
        return (int32_t)v;
    }
}

template <typename FT, bool SWAP>
__global__ __launch_bounds__(256) void k_f32_max(const FT *__restrict__ raw, long long n, float *__restrict__ part)
{
    float m = 0.0f;   // gft starts its maximum at 0.0 (gft_scene32.cpp:577-583)
    for (long long i = gtid(); i < n; i += gstride()) {
        const float v = (float)load_sample<FT, SWAP>(raw, i);
        if (m < v) m = v;
    }
    m = block_reduce(m, OpMax());
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// convert into int32 scratch and reduce min/max into per-workgroup slots
template <typename FT, bool SWAP>
__global__ __launch_bounds__(256) void k_convert(const FT *__restrict__ raw, long long n, const float *__restrict__ fmax,
                                                 int npart, int32_t *__restrict__ out, int *__restrict__ mm)
{
    __shared__ float s_factor;
    if (threadIdx.x == 0) {   // fold the workgroup maxima of k_f32_max (max is order-free)
        float m = 0.0f;
        if (fmax)
            for (int i = 0; i < npart; i++) m = fmaxf(m, fmax[i]);
        s_factor = fmax ? (float)(10000.0 / (double)m) : 1.0f;
    }
    __syncthreads();
    const float factor = s_factor;
    int lo = INT_MAX, hi = INT_MIN;
    for (long long i = gtid(); i < n; i += gstride()) {
        const int32_t v = to_i32<FT, SWAP>(raw, i, factor);
        out[i] = v;
        lo = min(lo, v);
        hi = max(hi, v);
    }
    lo = block_reduce(lo, OpMin());
    hi = block_reduce(hi, OpMax());
    if (threadIdx.x == 0) {
        mm[2 * blockIdx.x] = lo;
        mm[2 * blockIdx.x + 1] = hi;
    }
}

// out = in + shift, as int16 or int32
template <typename DT>
__global__ __launch_bounds__(256) void k_shift_store(const int32_t *__restrict__ in, long long n, int shift,
                                                     DT *__restrict__ out)
{
    for (long long i = gtid(); i < n; i += gstride()) out[i] = (DT)(in[i] + shift);
}

void pinned_ensure(roift_ctx &c, size_t bytes)
{
    if (c.pinned_bytes >= bytes) return;
    for (auto &p : c.pinned)
        if (p) {
            ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
            ROIFT_HIP_CHECK(hipHostFree(p));
            p = nullptr;
        }
    for (auto &p : c.pinned)
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) io_fail(ROIFT_ENOMEM, "hipHostMalloc failed");
    c.pinned_bytes = bytes;
    for (auto &e : c.io_ev)
        if (!e) ROIFT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

constexpr size_t IO_CHUNK = 8u << 20;

}  // namespace

// ---- host entry points -------------------------------------------------------------

const char *io_last_error() { return g_io_error.c_str(); }

int volume_info(const char *path, roift_volume_info *info)
{
    return io_guard([&]() -> int {
        if (!path || !info) io_fail(ROIFT_EINVAL, "null argument");
        Header h;
        read_header(path, h);
        *info = roift_volume_info{h.xs, h.ys, h.zs, h.dx, h.dy, h.dz, h.stored, h.format};
        return ROIFT_OK;
    });
}

int volume_read(const char *path, int32_t *out, long long n)
{
    return io_guard([&]() -> int {
        if (!path || !out) io_fail(ROIFT_EINVAL, "null argument");
        Header h;
        read_header(path, h);
        if (n != h.n()) io_fail(ROIFT_EINVAL, "buffer size does not match the volume");
        std::vector<unsigned char> raw((size_t)h.n() * h.esize);
        InFile f(h.data_path);
        f.skip_to(h.data_off);
        f.read(raw.data(), raw.size(), "the samples");
        if (h.stored == ROIFT_F32) {
            float mx = 0.0f;
            for (long long i = 0; i < n; i++) {
                const float v = sample_f32(&raw[(size_t)i * 4], h.swap);
                if (mx < v) mx = v;
            }
            const float factor = f32_factor(mx);
            for (long long i = 0; i < n; i++) out[i] = gft_round_f(sample_f32(&raw[(size_t)i * 4], h.swap) * factor);
        } else {
            for (long long i = 0; i < n; i++) out[i] = sample(&raw[(size_t)i * h.esize], h.stored, h.swap);
        }
        return ROIFT_OK;
    });
}

int volume_write(const char *path, const int32_t *data, const roift_volume_info *info, const char *like)
{
    return io_guard([&]() -> int {
        if (!path || !data || !info) io_fail(ROIFT_EINVAL, "null argument");
        const std::string p(path);
        const long long n = (long long)info->xsize * info->ysize * info->zsize;
        if (info->xsize <= 0 || info->ysize <= 0 || info->zsize <= 0) io_fail(ROIFT_EINVAL, "bad volume size");
        if (ends_with(p, ".scn")) {
            FILE *f = std::fopen(path, "wb");
            if (!f) io_fail(ROIFT_EINVAL, "cannot create " + p);
            int mx = INT_MIN;
            for (long long i = 0; i < n; i++) mx = std::max(mx, data[i]);
            std::fprintf(f, "SCN\n%d %d %d\n%f %f %f\n", info->xsize, info->ysize, info->zsize, info->dx, info->dy,
                         info->dz);
            bool ok = true;
            if (mx < 256) {   // samples cast to the narrower unsigned type, as gft writes them
                std::fprintf(f, "8\n");
                std::vector<uint8_t> b((size_t)n);
                for (long long i = 0; i < n; i++) b[(size_t)i] = (uint8_t)data[i];
                ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
            } else if (mx < 65536) {
                std::fprintf(f, "16\n");
                std::vector<uint16_t> b((size_t)n);
                for (long long i = 0; i < n; i++) b[(size_t)i] = (uint16_t)data[i];
                ok = std::fwrite(b.data(), 2, b.size(), f) == b.size();
            } else {
                std::fprintf(f, "32\n");
                ok = std::fwrite(data, 4, (size_t)n, f) == (size_t)n;
            }
            if (std::fclose(f) != 0 || !ok) io_fail(ROIFT_EINVAL, "write failed: " + p);
            return ROIFT_OK;
        }
        if (!(ends_with(p, ".nii") || ends_with(p, ".nii.gz")))
            io_fail(ROIFT_EUNSUPPORTED, p + ": labels are written as .scn, .nii or .nii.gz");
        unsigned char h[352] = {};
        Header src;
        if (like && *like) {
            read_header(like, src);
            if (src.format == 2) std::memcpy(h, src.nii, 348);
        }
        if (src.format != 2) {   // a fresh header: identity orientation, millimetres
            hset<float>(h, 76, 1.0f);          // pixdim[0] = qfac
            hset<char>(h, 123, (char)2);       // xyzt_units: mm
        }
        hset<int>(h, 0, 348);
        const short dim[8] = {3, (short)info->xsize, (short)info->ysize, (short)info->zsize, 1, 1, 1, 1};
        if (info->xsize > 32767 || info->ysize > 32767 || info->zsize > 32767) io_fail(ROIFT_EINVAL, "dims exceed NIfTI-1");
        for (int i = 0; i < 8; i++) hset<short>(h, 40 + 2 * i, dim[i]);
        hset<short>(h, 70, (short)NII_INT32);
        hset<short>(h, 72, (short)32);
        hset<float>(h, 80, info->dx);
        hset<float>(h, 84, info->dy);
        hset<float>(h, 88, info->dz);
        hset<float>(h, 108, 352.0f);
        std::memcpy(h + 344, "n+1", 4);
        const bool gz = ends_with(p, ".gz");
        gzFile f = gzopen(path, gz ? "wb6" : "wbT");
        if (!f) io_fail(ROIFT_EINVAL, "cannot create " + p);
        bool ok = gzwrite(f, h, 352) == 352;
        const char *d = (const char *)data;
        size_t left = (size_t)n * 4;
        while (ok && left > 0) {
            const unsigned chunk = (unsigned)std::min<size_t>(left, 1u << 30);
            ok = gzwrite(f, d, chunk) == (int)chunk;
            d += chunk;
            left -= chunk;
        }
        if (gzclose(f) != Z_OK || !ok) io_fail(ROIFT_EINVAL, "write failed: " + p);
        return ROIFT_OK;
    });
}

int seeds_read(const char *path, int xs, int ys, int zs, int32_t *seeds, int32_t *labels, int cap, int *count,
               int *clamped)
{
    return io_guard([&]() -> int {
        if (!path || !count) io_fail(ROIFT_EINVAL, "null argument");
        FILE *f = std::fopen(path, "r");
        if (!f) io_fail(ROIFT_EINVAL, std::string("cannot open seeds file ") + path);
        int n = 0, k = 0, neg = 0;
        if (std::fscanf(f, " %d", &n) != 1 || n < 0) {
            std::fclose(f);
            io_fail(ROIFT_EINVAL, "malformed seeds file");
        }
        for (int i = 0; i < n; i++) {
            int x, y, z, id, lb;
            if (std::fscanf(f, " %d %d %d %d %d", &x, &y, &z, &id, &lb) != 5) break;   // as fscanf leaves them
            if (lb < 0) {
                lb = 0;
                neg = 1;
            }
            if (x < 0 || x >= xs || y < 0 || y >= ys || z < 0 || z >= zs) continue;
            if (k < cap && seeds) seeds[k] = x + y * xs + z * xs * ys;
            if (k < cap && labels) labels[k] = lb;
            k++;
        }
        std::fclose(f);
        *count = k;
        if (clamped) *clamped = neg;
        if ((seeds || labels) && k > cap) io_fail(ROIFT_EINVAL, "more seeds than the buffer holds");
        return ROIFT_OK;
    });
}

int polarity_read(const char *path, float *per_class, int cap, int *count)
{
    return io_guard([&]() -> int {
        if (!path || !count) io_fail(ROIFT_EINVAL, "null argument");
        *count = 0;
        FILE *f = std::fopen(path, "r");
        if (!f) io_fail(ROIFT_EINVAL, std::string("cannot open polarity file ") + path);
        int n = 0;
        std::vector<float> v;
        bool ok = std::fscanf(f, " %d", &n) == 1 && n > 0;
        for (int i = 0; ok && i < n; i++) {
            float pv;
            ok = std::fscanf(f, " %f", &pv) == 1;
            v.push_back(pv * 100.0f);
        }
        std::fclose(f);
        if (!ok) io_fail(ROIFT_EINVAL, "malformed polarity file");
        if (n > cap) io_fail(ROIFT_EINVAL, "more classes than the buffer holds");
        std::copy(v.begin(), v.end(), per_class);
        *count = n;
        return ROIFT_OK;
    });
}

int volume_read_dev(roift_ctx &c, const char *path, void *d_dst, int dst_dtype, int flags, int *h_shift,
                    roift_volume_info *info)
{
    int rc = io_guard([&]() -> int {
        if (!path || !d_dst) io_fail(ROIFT_EINVAL, "null argument");
        if (dst_dtype != ROIFT_I16 && dst_dtype != ROIFT_I32) io_fail(ROIFT_EINVAL, "destination must be I16 or I32");
        Header h;
        read_header(path, h);
        if (info) *info = roift_volume_info{h.xs, h.ys, h.zs, h.dx, h.dy, h.dz, h.stored, h.format};
        const long long n = h.n();
        const size_t bytes = (size_t)n * h.esize;
        char *raw = buf<char>(c, "io.raw", bytes);
        pinned_ensure(c, IO_CHUNK);
        InFile f(h.data_path);
        f.skip_to(h.data_off);
        // double-buffered: read chunk k into pinned[k & 1] while chunk k-1 is being copied
        int k = 0;
        for (size_t off = 0; off < bytes; off += IO_CHUNK, k++) {
            const size_t len = std::min(IO_CHUNK, bytes - off);
            const int b = k & 1;
            if (k >= 2) ROIFT_HIP_CHECK(hipEventSynchronize(c.io_ev[b]));
            f.read(c.pinned[b], len, "the samples");
            ROIFT_HIP_CHECK(hipMemcpyAsync(raw + off, c.pinned[b], len, hipMemcpyHostToDevice, c.stream));
            ROIFT_HIP_CHECK(hipEventRecord(c.io_ev[b], c.stream));
        }
        const unsigned G = grid_for(n, 256, 1024);
        int32_t *tmp = buf<int32_t>(c, "io.i32", (size_t)n);
        int *mm = buf<int>(c, "io.mm", 2 * (size_t)G);
        float *fpart = nullptr;
        auto launch = [&](auto tag, auto swaptag) {
            using FT = decltype(tag);
            constexpr bool SW = decltype(swaptag)::value;
            if (h.stored == ROIFT_F32) {
                fpart = buf<float>(c, "io.fmax", G);
                hipLaunchKernelGGL((k_f32_max<FT, SW>), dim3(G), dim3(256), 0, c.stream, (const FT *)raw, n, fpart);
                ROIFT_HIP_CHECK(hipGetLastError());
            }
            hipLaunchKernelGGL((k_convert<FT, SW>), dim3(G), dim3(256), 0, c.stream, (const FT *)raw, n,
                               (const float *)fpart, (int)G, tmp, mm);
            ROIFT_HIP_CHECK(hipGetLastError());
        };
        using T = std::true_type;
        using F = std::false_type;
        switch (h.stored) {
        case ROIFT_U8: launch(uint8_t{}, F{}); break;
        case ROIFT_I8: launch(int8_t{}, F{}); break;
        case ROIFT_U16: h.swap ? launch(uint16_t{}, T{}) : launch(uint16_t{}, F{}); break;
        case ROIFT_I16: h.swap ? launch(int16_t{}, T{}) : launch(int16_t{}, F{}); break;
        case ROIFT_I32: h.swap ? launch(int32_t{}, T{}) : launch(int32_t{}, F{}); break;
        default: h.swap ? launch(float{}, T{}) : launch(float{}, F{}); break;
        }
        std::vector<int> hm(2 * (size_t)G);
        ROIFT_HIP_CHECK(hipMemcpyAsync(hm.data(), mm, hm.size() * sizeof(int), hipMemcpyDeviceToHost, c.stream));
        ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
        long long lo = INT_MAX, hi = INT_MIN;
        for (unsigned b = 0; b < G; b++) {
            lo = std::min<long long>(lo, hm[2 * b]);
            hi = std::max<long long>(hi, hm[2 * b + 1]);
        }
        // the CLI's shift (oiftrelax.cpp:263-268): a negative minimum is added back
        const int shift = ((flags & ROIFT_READ_SHIFT_MIN) && lo < 0) ? (int)-lo : 0;
        if (h_shift) *h_shift = shift;
        if (hi + shift > INT_MAX) io_fail(ROIFT_EINVAL, "shifted intensities overflow int32");
        if (dst_dtype == ROIFT_I16 && (lo + shift < -32768 || hi + shift > 32767))
            io_fail(ROIFT_EINVAL, "intensities do not fit int16 (read as I32)");
        if (dst_dtype == ROIFT_I16)
            hipLaunchKernelGGL(k_shift_store<int16_t>, dim3(G), dim3(256), 0, c.stream, (const int32_t *)tmp, n, shift,
                               (int16_t *)d_dst);
        else
            hipLaunchKernelGGL(k_shift_store<int32_t>, dim3(G), dim3(256), 0, c.stream, (const int32_t *)tmp, n, shift,
                               (int32_t *)d_dst);
        ROIFT_HIP_CHECK(hipGetLastError());
        ROIFT_HIP_CHECK(hipStreamSynchronize(c.stream));
        return ROIFT_OK;
    });
    if (rc != ROIFT_OK) c.last_error = g_io_error;
    return rc;
}

}  // namespace roift
