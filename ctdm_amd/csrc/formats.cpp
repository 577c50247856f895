// The reference's on-disk formats, host side (no device work, no context):
//   .phimap  one retrieved frame (outputImg / openImage, HilbertTransform.cu:2437-2477): a 1-byte status,
//            the two illumination angles as doubles, sizeX and sizeY as 32-bit ints, then sizeX * sizeY
//            (phase, amplitude) float pairs (struct microImg, HilbertTransform.cuh:46-50); little endian,
//            fields written one after another without padding.
//   .rimap   the reconstructed volume (BatchRecon export, HilbertTransform.cu:2013-2041; exportComplex,
//            exportRAW_F :3457-3487): the raw V^3 complex64 array, or only its real or imaginary floats.
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "internal.hpp"

namespace {

struct File {
    std::FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

template <class T>
bool put(std::FILE* f, const T& v)
{
    return std::fwrite(&v, sizeof(T), 1, f) == 1;
}
template <class T>
bool get(std::FILE* f, T& v)
{
    return std::fread(&v, sizeof(T), 1, f) == 1;
}

}  // namespace

extern "C" {

ctdm_status ctdm_phimap_write(const char* path, int status, double ang_x, double ang_y, int sizeX, int sizeY,
                              const float* phase, const float* amp)
{
    if (!path || !phase || !amp || sizeX <= 0 || sizeY <= 0) return CTDM_EINVAL;
    File fl(path, "wb");
    if (!fl.f) return CTDM_EINVAL;
    const uint8_t st = status ? 1 : 0;
    bool ok = put(fl.f, st) && put(fl.f, ang_x) && put(fl.f, ang_y) && put(fl.f, (int32_t)sizeX) &&
              put(fl.f, (int32_t)sizeY);
    const size_t n = (size_t)sizeX * sizeY;
    std::vector<float> buf(2 * std::min<size_t>(n, 1 << 20));
    for (size_t i0 = 0; ok && i0 < n; i0 += buf.size() / 2) {
        const size_t m = std::min(n - i0, buf.size() / 2);
        for (size_t k = 0; k < m; ++k) {
            buf[2 * k] = phase[i0 + k];
            buf[2 * k + 1] = amp[i0 + k];
        }
        ok = std::fwrite(buf.data(), sizeof(float), 2 * m, fl.f) == 2 * m;
    }
    return ok ? CTDM_OK : CTDM_EDEVICE;
}

ctdm_status ctdm_phimap_read(const char* path, int* status, double* ang_x, double* ang_y, int* sizeX, int* sizeY,
                             float* phase, float* amp, size_t capacity)
{
    if (!path) return CTDM_EINVAL;
    File fl(path, "rb");
    if (!fl.f) return CTDM_EINVAL;
    uint8_t st = 0;
    double ax = 0, ay = 0;
    int32_t sx = 0, sy = 0;
    if (!(get(fl.f, st) && get(fl.f, ax) && get(fl.f, ay) && get(fl.f, sx) && get(fl.f, sy))) return CTDM_EINVAL;
    if (sx <= 0 || sy <= 0) return CTDM_EINVAL;
    if (status) *status = st ? 1 : 0;
    if (ang_x) *ang_x = ax;
    if (ang_y) *ang_y = ay;
    if (sizeX) *sizeX = sx;
    if (sizeY) *sizeY = sy;
    if (!phase && !amp) return CTDM_OK;  // header only
    const size_t n = (size_t)sx * sy;
    if (!phase || !amp || capacity < n) return CTDM_EINVAL;
    std::vector<float> buf(2 * std::min<size_t>(n, 1 << 20));
    for (size_t i0 = 0; i0 < n; i0 += buf.size() / 2) {
        const size_t m = std::min(n - i0, buf.size() / 2);
        if (std::fread(buf.data(), sizeof(float), 2 * m, fl.f) != 2 * m) return CTDM_EINVAL;
        for (size_t k = 0; k < m; ++k) {
            phase[i0 + k] = buf[2 * k];
            amp[i0 + k] = buf[2 * k + 1];
        }
    }
    return CTDM_OK;
}

ctdm_status ctdm_rimap_write(const char* path, const ctdm_complex* ri, size_t n, int part)
{
    if (!path || !ri || part < 0 || part > 2) return CTDM_EINVAL;
    File fl(path, "wb");
    if (!fl.f) return CTDM_EINVAL;
    if (part == 0) return std::fwrite(ri, sizeof(ctdm_complex), n, fl.f) == n ? CTDM_OK : CTDM_EDEVICE;
    std::vector<float> buf(std::min<size_t>(n, 1 << 20));
    for (size_t i0 = 0; i0 < n; i0 += buf.size()) {
        const size_t m = std::min(n - i0, buf.size());
        for (size_t k = 0; k < m; ++k) buf[k] = part == 1 ? ri[i0 + k].x : ri[i0 + k].y;
        if (std::fwrite(buf.data(), sizeof(float), m, fl.f) != m) return CTDM_EDEVICE;
    }
    return CTDM_OK;
}

ctdm_status ctdm_rimap_read(const char* path, ctdm_complex* ri, size_t n, int part)
{
    if (!path || !ri || part < 0 || part > 2) return CTDM_EINVAL;
    File fl(path, "rb");
    if (!fl.f) return CTDM_EINVAL;
    if (part == 0) return std::fread(ri, sizeof(ctdm_complex), n, fl.f) == n ? CTDM_OK : CTDM_EINVAL;
    std::vector<float> buf(std::min<size_t>(n, 1 << 20));
    for (size_t i0 = 0; i0 < n; i0 += buf.size()) {
        const size_t m = std::min(n - i0, buf.size());
        if (std::fread(buf.data(), sizeof(float), m, fl.f) != m) return CTDM_EINVAL;
        for (size_t k = 0; k < m; ++k) {
            if (part == 1)
                ri[i0 + k].x = buf[k];
            else
                ri[i0 + k].y = buf[k];
        }
    }
    return CTDM_OK;
}

}  // extern "C"
