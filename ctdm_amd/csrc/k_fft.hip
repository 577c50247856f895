// Generic batched 1-D FFT passes over pencils of strided arrays, plus 2-D / 3-D quadrant shifts.
//
// These are the building blocks that replace the reference's cuFFT calls (FFT.cu:636-838) on the
// plain, reference-structured path: a 2-D transform is a row pass plus a column pass, a 3-D
// transform is an x, a y and a z pass. The fused kernels of the batched path reuse the same device
// transforms (fft_dev.hpp) inside their own prologues and epilogues.
#include "fft_dev.hpp"
#include "internal.hpp"

namespace ctdm {

// One workgroup transforms B pencils. Pencil b of tile `tin` in plane `pl` starts at
//   data + pl*stride_plane + (tin*B + b)*stride_b   and has elements at multiples of stride_n.
// BFAST: consecutive lanes walk b (coalesced when stride_b == 1); otherwise lanes walk the
// transform's own threads (coalesced when stride_n == 1).
template <int N, int B, bool BFAST, int DIR>
__global__ __launch_bounds__(B* Plan<N>::T) void k_fft_pencils(
    float2* __restrict__ data, long long stride_n, long long stride_b, int tiles_per_plane,
    long long stride_plane, long long npencils_per_plane, const float2* __restrict__ tw, float scale,
    int conj_out)
{
    constexpr int T = Plan<N>::T;
    constexpr int LS = LdsStride<N, BFAST ? B : 1>::value;
    __shared__ float2 lds[B * LS];
    __shared__ float2 twl[LdsSize<N>::value];
    stage_twiddles<N>(twl, tw);
    const TwPad twp{twl};
    const int tid = threadIdx.x;
    const int b = BFAST ? (tid % B) : (tid / T);
    const int t = BFAST ? (tid / B) : (tid % T);
    const long long pl = blockIdx.x / tiles_per_plane;
    const long long tin = blockIdx.x % tiles_per_plane;
    const long long pencil = tin * B + b;
    const bool valid = pencil < npencils_per_plane;
    float2* base = data + pl * stride_plane + (valid ? pencil : 0) * stride_b;
    float2* s = lds + b * LS;
    auto gl = [&](int i) -> float2 { return valid ? base[(long long)i * stride_n] : make_float2(0.f, 0.f); };
    auto gs = [&](int i, float2 v) {
        if (!valid) return;
        v = cscale(v, scale);
        if (conj_out) v.y = -v.y;
        base[(long long)i * stride_n] = v;
    };
    Plan<N>::template run<DIR>(t, s, twp, gl, gs);
}

template <int N, int B, bool BFAST>
static void launch_pencils(Ctx& c, float2* data, long long stride_n, long long stride_b,
                           long long npencils, long long nplanes, long long stride_plane, int dir,
                           float scale, bool conj_out)
{
    const int tiles = (int)((npencils + B - 1) / B);
    const long long blocks = (long long)tiles * nplanes;
    CTDM_REQUIRE(blocks < (1LL << 31), "fft pass: too many pencils");
    dim3 grid((unsigned)blocks), block(B * Plan<N>::T);
    const float2* tw = c.tw(N);
    if (dir < 0)
        hipLaunchKernelGGL((k_fft_pencils<N, B, BFAST, -1>), grid, block, 0, c.stream, data, stride_n,
                           stride_b, tiles, stride_plane, npencils, tw, scale, (int)conj_out);
    else
        hipLaunchKernelGGL((k_fft_pencils<N, B, BFAST, +1>), grid, block, 0, c.stream, data, stride_n,
                           stride_b, tiles, stride_plane, npencils, tw, scale, (int)conj_out);
    CTDM_CHECK_LAUNCH("k_fft_pencils");
}

// Pencils per workgroup so that B * T = 256 threads.
template <int N>
struct RowB {
    static constexpr int value = 256 / Plan<N>::T;
};

template <bool BFAST>
static void dispatch(Ctx& c, int n, float2* data, long long stride_n, long long stride_b,
                     long long npencils, long long nplanes, long long stride_plane, int dir,
                     float scale, bool conj_out)
{
    switch (n) {
    case 32: launch_pencils<32, RowB<32>::value, BFAST>(c, data, stride_n, stride_b, npencils, nplanes, stride_plane, dir, scale, conj_out); break;
    case 64: launch_pencils<64, RowB<64>::value, BFAST>(c, data, stride_n, stride_b, npencils, nplanes, stride_plane, dir, scale, conj_out); break;
    case 128: launch_pencils<128, RowB<128>::value, BFAST>(c, data, stride_n, stride_b, npencils, nplanes, stride_plane, dir, scale, conj_out); break;
    case 192: launch_pencils<192, RowB<192>::value, BFAST>(c, data, stride_n, stride_b, npencils, nplanes, stride_plane, dir, scale, conj_out); break;
    case 256: launch_pencils<256, RowB<256>::value, BFAST>(c, data, stride_n, stride_b, npencils, nplanes, stride_plane, dir, scale, conj_out); break;
    case 384: launch_pencils<384, RowB<384>::value, BFAST>(c, data, stride_n, stride_b, npencils, nplanes, stride_plane, dir, scale, conj_out); break;
    case 512: launch_pencils<512, RowB<512>::value, BFAST>(c, data, stride_n, stride_b, npencils, nplanes, stride_plane, dir, scale, conj_out); break;
    case 1024: launch_pencils<1024, RowB<1024>::value, BFAST>(c, data, stride_n, stride_b, npencils, nplanes, stride_plane, dir, scale, conj_out); break;
    default: fail(CTDM_EUNSUPPORTED, "fft length " + std::to_string(n) + " has no kernel");
    }
}

// `count` contiguous rows of length n.
void fft_rows(Ctx& c, float2* data, int n, long long count, int dir, float scale, bool conj_out)
{
    dispatch<false>(c, n, data, 1, n, count, 1, 0, dir, scale, conj_out);
}

// Columns of length n: element (x, i, plane) at data + plane*stride_plane + x + i*stride_n,
// x < width.
void fft_cols(Ctx& c, float2* data, int n, int width, long long stride_n, long long nplanes,
              long long stride_plane, int dir, float scale, bool conj_out)
{
    dispatch<true>(c, n, data, stride_n, 1, width, nplanes, stride_plane, dir, scale, conj_out);
}

// ---------------------------------------------------------------- shifts
// In-place quadrant swap of even-sized W x H images (fftshift == ifftshift), reference
// cuFFT2Dshift (FFT.cu:220-236).
__global__ void k_fftshift2d(float2* __restrict__ d, int W, int H, long long batch)
{
    const long long n = (long long)(W / 2) * (H / 2);
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * batch) return;
    const long long img = g / n;
    const int r = (int)(g % n);
    const int i = r % (W / 2), j = r / (W / 2);
    float2* p = d + img * (long long)W * H;
    const long long a = i + (long long)W * j, bq = (i + W / 2) + (long long)W * (j + H / 2);
    const long long c2 = (i + W / 2) + (long long)W * j, d2 = i + (long long)W * (j + H / 2);
    float2 t = p[a];
    p[a] = p[bq];
    p[bq] = t;
    t = p[c2];
    p[c2] = p[d2];
    p[d2] = t;
}

void fftshift2d(Ctx& c, float2* data, int W, int H, long long batch)
{
    CTDM_REQUIRE(W % 2 == 0 && H % 2 == 0, "fftshift2d: even sizes required");
    const long long n = (long long)(W / 2) * (H / 2) * batch;
    hipLaunchKernelGGL(k_fftshift2d, dim3(grid1d(n, 256)), dim3(256), 0, c.stream, data, W, H, batch);
    CTDM_CHECK_LAUNCH("k_fftshift2d");
}

// In-place octant swap of a V^3 volume, reference cuFFT3Dshift (FFT.cu:324-350).
__global__ void k_fftshift3d(float2* __restrict__ d, int V)
{
    const int h = V / 2;
    const long long n = (long long)h * h * h;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int i = (int)(g % h), j = (int)((g / h) % h), k = (int)(g / ((long long)h * h));
    const long long VV = (long long)V * V;
    auto at = [&](int x, int y, int z) -> float2& { return d[x + (long long)V * y + VV * z]; };
    float2 t;
    t = at(i, j, k); at(i, j, k) = at(i + h, j + h, k + h); at(i + h, j + h, k + h) = t;
    t = at(i, j + h, k); at(i, j + h, k) = at(i + h, j, k + h); at(i + h, j, k + h) = t;
    t = at(i + h, j, k); at(i + h, j, k) = at(i, j + h, k + h); at(i, j + h, k + h) = t;
    t = at(i + h, j + h, k); at(i + h, j + h, k) = at(i, j, k + h); at(i, j, k + h) = t;
}

void fftshift3d(Ctx& c, float2* data, int V)
{
    CTDM_REQUIRE(V % 2 == 0, "fftshift3d: even size required");
    const long long n = (long long)(V / 2) * (V / 2) * (V / 2);
    hipLaunchKernelGGL(k_fftshift3d, dim3(grid1d(n, 256)), dim3(256), 0, c.stream, data, V);
    CTDM_CHECK_LAUNCH("k_fftshift3d");
}

}  // namespace ctdm
