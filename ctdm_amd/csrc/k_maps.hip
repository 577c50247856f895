// Map sanitising before gridding: the 3x3 median of BatchRecon's phase and amplitude stacks
// (MedianFilter / MedianFilter_gpu, Recon_kernel.cu:7-27, 65-119) and checkMaps (Recon_kernel.cu:121-142),
// followed by the logarithm of cuCopyArray (Recon_kernel.cu:154-169).
//
// The reference filters each map in place while neighbouring threads overwrite their pixels, so its
// result depends on thread order. Here the filter is out of place: every output pixel is the median of
// the unfiltered input. Within that definition the selection is the reference's: the nine values are
// gathered column by column (x - 1, x, x + 1 outer; y - 1, y, y + 1 inner) and the fifth pick of its
// selection sort (strict '<', first minimum kept) is the result, so NaN inputs resolve the same way.
// Pixels in the first and last row and column are copied.
#include "internal.hpp"

namespace ctdm {

__device__ __forceinline__ float median9(const float* __restrict__ m, int W, int x, int y)
{
    float v[9];
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) v[(dx + 1) * 3 + (dy + 1)] = m[(long long)(y + dy) * W + (x + dx)];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        int mn = i;
#pragma unroll
        for (int l = i + 1; l < 9; ++l)
            if (v[l] < v[mn]) mn = l;
        const float t = v[i];
        v[i] = v[mn];
        v[mn] = t;
    }
    return v[4];
}

__device__ __forceinline__ float filtered(const float* __restrict__ m, int W, int H, int x, int y, bool median)
{
    if (median && x > 0 && x < W - 1 && y > 0 && y < H - 1) return median9(m, W, x, y);
    return m[(long long)y * W + x];
}

__global__ void k_median3x3(const float* __restrict__ in, float* __restrict__ out, int batch, int H, int W)
{
    const long long n = (long long)H * W;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * batch) return;
    const long long b = g / n;
    const int p = (int)(g % n), x = p % W, y = p / W;
    out[g] = filtered(in + b * n, W, H, x, y, true);
}

// checkMaps: phase NaN/Inf -> 0; amplitude NaN/Inf -> 1, 0 -> 0.01, otherwise |A|.
__device__ __forceinline__ float check_phase(float p) { return isfinite(p) ? p : 0.f; }
__device__ __forceinline__ float check_amp(float a)
{
    if (!isfinite(a)) return 1.f;
    if (a == 0.f) return 0.01f;
    return fabsf(a);
}

__global__ void k_check_maps(float* __restrict__ phase, float* __restrict__ amp, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (phase) phase[i] = check_phase(phase[i]);
    if (amp) amp[i] = check_amp(amp[i]);
}

// Median (optional), checkMaps and the logarithm in one pass: phase and second map in, unwrapped phase
// and ln A out, ready for the Rytov spectrum. amplitude = 0: the second map already holds L = ln A;
// the median commutes with the monotone logarithm, and checkMaps on A = exp(L) becomes
// L = NaN or +inf -> 0, L = -inf -> ln 0.01.
__global__ void k_prepare_maps(const float* __restrict__ ph, const float* __restrict__ m2, float* __restrict__ ph_out,
                               float* __restrict__ la_out, int batch, int M, int median, int amplitude)
{
    const long long n = (long long)M * M;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * batch) return;
    const long long b = g / n;
    const int p = (int)(g % n), x = p % M, y = p / M;
    ph_out[g] = check_phase(filtered(ph + b * n, M, M, x, y, median));
    const float v = filtered(m2 + b * n, M, M, x, y, median);
    float la;
    if (amplitude) {
        la = logf(check_amp(v));
    } else {
        la = (isnan(v) || v == INFINITY) ? 0.f : (v == -INFINITY ? logf(0.01f) : v);
    }
    la_out[g] = la;
}

void median3x3(Ctx& c, const float* in, float* out, int batch, int H, int W)
{
    CTDM_REQUIRE(in && out && in != out, "median3x3: null or aliased pointer");
    CTDM_REQUIRE(batch >= 0 && H > 0 && W > 0, "median3x3: bad sizes");
    const long long n = (long long)batch * H * W;
    if (n == 0) return;
    hipLaunchKernelGGL(k_median3x3, dim3(grid1d(n, 256)), dim3(256), 0, c.stream, in, out, batch, H, W);
    CTDM_CHECK_LAUNCH("k_median3x3");
}

void check_maps(Ctx& c, float* phase, float* amp, long long n)
{
    CTDM_REQUIRE(n >= 0 && (phase || amp || n == 0), "check_maps: null pointers");
    if (n == 0) return;
    hipLaunchKernelGGL(k_check_maps, dim3(grid1d(n, 256)), dim3(256), 0, c.stream, phase, amp, n);
    CTDM_CHECK_LAUNCH("k_check_maps");
}

void prepare_maps(Ctx& c, const float* ph, const float* m2, float* ph_out, float* la_out, int batch, int M,
                  unsigned flags)
{
    const long long n = (long long)batch * M * M;
    if (n == 0) return;
    hipLaunchKernelGGL(k_prepare_maps, dim3(grid1d(n, 256)), dim3(256), 0, c.stream, ph, m2, ph_out, la_out, batch,
                       M, (int)((flags & CTDM_MAPS_MEDIAN) != 0), (int)((flags & CTDM_MAPS_AMPLITUDE) != 0));
    CTDM_CHECK_LAUNCH("k_prepare_maps");
}

}  // namespace ctdm
