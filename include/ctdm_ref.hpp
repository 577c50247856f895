// ctdm_ref.hpp — C++ host layer over the libctdm C ABI that keeps the reference program's function
// names and argument meaning, so the reference's per-frame loop (HilbertTransform.cu:785-1165) and
// BatchRecon call (:1243) map one to one. Differences from the reference, all deliberate:
//   - every pointer is device memory owned by the caller (the reference mixes host inputs and
//     driver-owned globals, HilbertTransform.cu:480-574, 1558-1559);
//   - errors throw ctdm_ref::Error instead of exit() (checkCudaErrors);
//   - the context carries the stream, scratch and FFT tables the reference kept in globals.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "ctdm.h"

namespace ctdm_ref {

struct Error : std::runtime_error {
    ctdm_status status;
    Error(ctdm_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

class Context {
public:
    explicit Context(int device = 0, ctdm_stream stream = nullptr)
    {
        ctdm_config cfg{};
        cfg.device = device;
        const ctdm_status s = ctdm_create(&ctx_, &cfg, stream);
        if (s != CTDM_OK) throw Error(s, "ctdm_create failed");
    }
    ~Context() { ctdm_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    ctdm_ctx* get() const { return ctx_; }
    void check(ctdm_status s) const
    {
        if (s != CTDM_OK) throw Error(s, ctdm_last_error(ctx_));
    }
    void sync() const { check(ctdm_sync(ctx_)); }

private:
    ctdm_ctx* ctx_ = nullptr;
};

// Pixel format of a frame pointer: float32, or the 8-bit pixels the reference reads from its BMP frames
// (unsigned char, HilbertTransform.cu:419-420, 2694) and 16-bit camera pixels, which the kernels widen as
// they load them (the reference widens on the host, :660-661).
inline ctdm_pixel pixel_of(const float*) { return CTDM_PIXEL_F32; }
inline ctdm_pixel pixel_of(const unsigned char*) { return CTDM_PIXEL_U8; }
inline ctdm_pixel pixel_of(const unsigned short*) { return CTDM_PIXEL_U16; }

// extractQPI_AlgorithmA(SP, BG, cuSP_FFT, cuBG_FFT, Nx, Ny, frame) (HilbertTransform.cu:1546)
// followed by calcWrapPhase (HilbertTransform.cu:3618): one Nx x Ny hologram pair in, the Nx/4
// (or crop.M, when set) wrapped phase and log-amplitude out. On frame 1 the sideband is located from
// BG and stored in `crop` for the following frames, as the reference does (HilbertTransform.cu:1572).
// P: float, unsigned char or unsigned short pixels.
template <class P>
inline void extractQPI_AlgorithmA(const Context& c, const P* SP, const P* BG, int Nx, int Ny, int frame,
                                  ctdm_crop& crop, float* wrapped, float* logamp)
{
    const ctdm_pixel px = pixel_of(SP);
    if (frame == 1 || crop.radius < 0) {
        const int M = crop.M > 0 ? crop.M : Nx / 4;
        c.check(ctdm_find_sideband_px(c.get(), BG, px, Ny, Nx, M, 0, &crop));
    }
    c.check(ctdm_retrieve_px(c.get(), SP, BG, px, 1, Ny, Nx, &crop, wrapped, logamp, nullptr));
}

// extractQPI_AlgorithmB(SP, BG, cuSP_FFT, cuBG_FFT, Nx, Ny, frame) (HilbertTransform.cu:1587) followed by
// calcWrapPhase: the reference's default retrieval. On frame 1 the sideband column is located from BG
// and kept in `center` for the following frames (sequence1DFFT, HilbertTransform.cu:1634). Outputs are
// (Nx/4) x (Ny/4) maps in the reference's transposed layout.
template <class P>
inline void extractQPI_AlgorithmB(const Context& c, const P* SP, const P* BG, int Nx, int Ny, int frame,
                                  int& center, float* wrapped, float* logamp)
{
    const ctdm_pixel px = pixel_of(SP);
    if (frame == 1 || center < 0) c.check(ctdm_find_sideband_rows_px(c.get(), BG, px, Ny, Nx, &center));
    c.check(ctdm_retrieve_rows_px(c.get(), SP, BG, px, 1, Ny, Nx, center, wrapped, logamp));
}

// AngCal_GPU(float *src, int sizeX, int sizeY, int frame, int totalFrame) (HilbertTransform.cu:2920) with
// the optics the reference reads from globals (dx = camera pixel [um], Mag, wavelength [nm], Nmed).
template <class P>
inline float AngCal_GPU(const Context& c, const P* src, int sizeX, int sizeY, double dx_um, double mag,
                        double wavelength_nm, double n_med)
{
    double a = 0;
    c.check(ctdm_angle_from_image_px(c.get(), src, pixel_of(src), 1, sizeY, sizeX, dx_um, mag, wavelength_nm, n_med,
                                      &a));
    return (float)a;
}

// DCT_UWLS_Unwrapped(float *ImgDst, float *ImgSrc, int sizeX, int sizeY) (DCT_Unwrapping.cu:10).
// flags = CTDM_UNWRAP_REFCOMPAT reproduces the reference's eigenvalues and IDCT exactly.
inline void DCT_UWLS_Unwrapped(const Context& c, float* ImgDst, const float* ImgSrc, int sizeX, int sizeY,
                               unsigned flags = CTDM_UNWRAP_WRAPDIFF)
{
    if (sizeX != sizeY) throw Error(CTDM_EINVAL, "DCT_UWLS_Unwrapped: square maps only");
    c.check(ctdm_unwrap_dct(c.get(), ImgSrc, ImgDst, 1, sizeX, flags));
}

// FastUnwrapping(float *ImgSrc, float *ImgDst, int sizeX, int sizeY) (FastUnwrapping.cu:5): the map
// rescaled to [0, 2 pi] (PhaseNormalize, :7), sin/cos-Laplacian right-hand side, Poisson solve,
// congruence rounding (ObtainUnwrap), minimum subtracted from the result (subtractMin, :108). The
// reference rescales ImgSrc itself in place; here the input is left untouched and the rescaled copy
// lives in the context's scratch.
inline void FastUnwrapping(const Context& c, const float* ImgSrc, float* ImgDst, int sizeX, int sizeY)
{
    if (sizeX != sizeY) throw Error(CTDM_EINVAL, "FastUnwrapping: square maps only");
    c.check(ctdm_unwrap_dct(c.get(), ImgSrc, ImgDst, 1, sizeX,
                            CTDM_UNWRAP_SINCOS | CTDM_UNWRAP_CONGRUENT | CTDM_UNWRAP_REFCOMPAT_FAST));
}

// cudaGoldsteinUnwrap2D(float *cu_phi, float *cu_UnWrapPhase, int sizeX, int sizeY, int frameNumber)
// (Goldstein.cu:5). The frame number only labelled the reference's debug output.
inline void cudaGoldsteinUnwrap2D(const Context& c, const float* cu_phi, float* cu_UnWrapPhase, int sizeX,
                                  int sizeY, int /*frameNumber*/)
{
    c.check(ctdm_goldstein(c.get(), cu_phi, 1, sizeY, sizeX, nullptr, nullptr, nullptr, cu_UnWrapPhase));
}

// BatchRecon(float *PhaseStack, float *AmpStack, float *sampleAngleRadX, float *sampleAngleRadY,
// int ReconFrame) (HilbertTransform.cu:1825): ReconFrame unwrapped-phase and amplitude maps of V x V
// (device), angles on the host, result = octant-shifted complex RI volume (device, V^3) that the
// reference wrote to a .rimap file (HilbertTransform.cu:1974-2041). As in the reference the maps are
// 3x3-median filtered and sanitised (MedianFilter, checkMaps) and the amplitude is logged on the device;
// pass recon_flags without CTDM_MAPS_MEDIAN to skip the filter, or without CTDM_MAPS_AMPLITUDE when the
// second stack already holds ln A.
inline void BatchRecon(const Context& c, const float* PhaseStack, const float* AmpStack, const float* sampleAngleRadX,
                       const float* sampleAngleRadY, int ReconFrame, int V, const ctdm_optics& optics, int iters,
                       ctdm_complex* ri, unsigned recon_flags = CTDM_MAPS_MEDIAN | CTDM_MAPS_AMPLITUDE)
{
    std::vector<ctdm_angle> ang(ReconFrame);
    for (int i = 0; i < ReconFrame; ++i) ang[i] = ctdm_angle{sampleAngleRadX[i], sampleAngleRadY[i]};
    c.check(ctdm_recon_from_maps(c.get(), PhaseStack, AmpStack, ang.data(), ReconFrame, V, &optics, iters,
                                 recon_flags, ri));
}

// MedianFilter(PhaseData, AmpData, Phase_img, Amp_img, nx, ny, nz) (Recon_kernel.cu:7-27) on device
// stacks: 3x3 median of both stacks into the outputs, then checkMaps on them.
inline void MedianFilter(const Context& c, float* PhaseData, float* AmpData, const float* Phase_img,
                         const float* Amp_img, int nx, int ny, int nz)
{
    c.check(ctdm_median3x3(c.get(), Phase_img, PhaseData, nz, ny, nx));
    c.check(ctdm_median3x3(c.get(), Amp_img, AmpData, nz, ny, nx));
    c.check(ctdm_check_maps(c.get(), PhaseData, AmpData, (size_t)nx * ny * nz));
}

}  // namespace ctdm_ref
