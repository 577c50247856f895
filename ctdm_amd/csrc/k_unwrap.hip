// Least-squares phase unwrapping by a DCT-II Poisson solve (reference DCT_UWLS_Unwrapped,
// DCT_Unwrapping.cu:10-107, 144-246, 300-312, 336-417), batched over maps, for gfx950.
//
// Maths (N = M, Neumann boundaries):
//   rho   = div of wrapped forward differences (NumericDerivative1 + SumDerivative), or with
//           CTDM_UNWRAP_SINCOS  cos(phi) L[sin phi] - sin(phi) L[cos phi] = sum_nbr sin(phi_n - phi)
//           (the trigonometric right-hand side of FastUnwrapping.cu:37-96 on the same stencil)
//   phi   = IDCT2( DCT2(rho) / lambda ),  lambda_ij = -4 (sin^2(pi i/2N) + sin^2(pi j/2N)), DC := 0
//           (REFCOMPAT_EIG: 2cos(4 pi i/(4N-1)) + 2cos(4 pi j/(4N-1)) - 4, DCT_Unwrapping.cu:310)
//   the division is done in FP64 (lambda -> 0 at low frequencies amplifies any rounding).
// Orthonormal DCT-II along a line via one N-point complex FFT (Makhoul reordering):
//   v[n] = x[2n], v[N-1-n] = x[2n+1];  X[k] = s_k Re(exp(-i pi k / 2N) FFT(v)[k]),
//   s_0 = sqrt(1/N), s_k = sqrt(2/N).
// Orthonormal DCT-III (its inverse):
//   V[k] = exp(i pi k/2N) (C_k - i C_{N-k}), C = X / s, C_N = 0;  v = IFFT(V) / N;
//   x[2n] = Re v[n], x[2n+1] = Re v[N-1-n].
// REFCOMPAT_IDCT reproduces the reference IDCT literally (2N Hermitian completion with the (i+1)
// twiddle on the right half, DCT_Unwrapping.cu:218-246), which needs a 2N-point transform.
//
// Batched path: three kernels per batch
//   1. k_uw_rows_fwd: RHS on the fly + DCT-II along x          -> D1[b][y][kx]
//   2. k_uw_cols:     DCT-II along y, / lambda, DCT-III along y -> D2[b][y][kx]
//   3. k_uw_rows_inv: DCT-III along x (+ congruence)           -> out[b][y][x]
#include "fft_dev.hpp"
#include "internal.hpp"

#include <cmath>
#include <mutex>

namespace ctdm {

enum : unsigned { F_SINCOS = CTDM_UNWRAP_SINCOS, F_REF_EIG = CTDM_UNWRAP_REFCOMPAT_EIG,
                  F_REF_IDCT = CTDM_UNWRAP_REFCOMPAT_IDCT, F_CONGRUENT = CTDM_UNWRAP_CONGRUENT,
                  F_REF_FAST = CTDM_UNWRAP_REFCOMPAT_FAST, F_ALL = 0x1fu };

static constexpr float kPI = 3.14159265358979323846f;

// Reference wrap of NumericDerivative1 (DCT_Unwrapping.cu:378-379).
__device__ __forceinline__ float wrapdiff(float t)
{
    const float s = (t > 0.f) ? 1.f : ((t < 0.f) ? -1.f : 0.f);
    return t - 2.f * kPI * s * floorf(fabsf(kPI * s + t) / (2.f * kPI));
}

// RHS at (i, j) of an N x N map.
__device__ __forceinline__ float rhs_at(const float* __restrict__ p, int N, int i, int j, bool sincos)
{
    const int k = i + N * j;
    const float c = p[k];
    if (sincos) {
        float r = 0.f;
        if (i < N - 1) r += sinf(p[k + 1] - c);
        if (i > 0) r += sinf(p[k - 1] - c);
        if (j < N - 1) r += sinf(p[k + N] - c);
        if (j > 0) r += sinf(p[k - N] - c);
        return r;
    }
    const float c1 = (i == N - 1) ? 0.f : wrapdiff(p[k + 1] - c);
    const float c2 = (i == 0) ? 0.f : wrapdiff(c - p[k - 1]);
    const float c3 = (j == N - 1) ? 0.f : wrapdiff(p[k + N] - c);
    const float c4 = (j == 0) ? 0.f : wrapdiff(c - p[k - N]);
    return (c1 - c2) + (c3 - c4);
}

// The same from rows staged in LDS: st holds map rows y0 - 1 .. y0 + G (row r of st is map row y0 - 1 + r;
// rows outside the map are never read, since the borders take no difference there).
__device__ __forceinline__ float rhs_staged(const float* st, int N, int i, int j, int y0, bool sincos)
{
    const float* row = st + (j - y0 + 1) * N;
    const float c = row[i];
    if (sincos) {
        float r = 0.f;
        if (i < N - 1) r += sinf(row[i + 1] - c);
        if (i > 0) r += sinf(row[i - 1] - c);
        if (j < N - 1) r += sinf(row[i + N] - c);
        if (j > 0) r += sinf(row[i - N] - c);
        return r;
    }
    const float c1 = (i == N - 1) ? 0.f : wrapdiff(row[i + 1] - c);
    const float c2 = (i == 0) ? 0.f : wrapdiff(c - row[i - 1]);
    const float c3 = (j == N - 1) ? 0.f : wrapdiff(row[i + N] - c);
    const float c4 = (j == 0) ? 0.f : wrapdiff(c - row[i - N]);
    return (c1 - c2) + (c3 - c4);
}

// Makhoul source index: position n of v holds x[src(n)].
__device__ __forceinline__ int mk_src(int n, int N) { return n < N / 2 ? 2 * n : 2 * (N - 1 - n) + 1; }

// ---------------------------------------------------------------- 1. RHS + DCT-II along x
// mode 0: input is a wrapped map, RHS computed on the fly; mode 1: input used as is (adjoint).
template <int N, int G>
__global__ __launch_bounds__(G* Plan<N>::T) void k_uw_rows_fwd(const float* __restrict__ in, int batch,
                                                              int mode, int sincos,
                                                              const float2* __restrict__ tw_n,
                                                              const float2* __restrict__ tw_4n,
                                                              float* __restrict__ out)
{
    constexpr int T = Plan<N>::T;
    __shared__ float2 lds[G * LdsSize<N>::value];
    const int g = threadIdx.x / T, t = threadIdx.x % T;
    const long long row = (long long)blockIdx.x * G + g;
    const bool valid = row < (long long)batch * N;
    const long long b = valid ? row / N : 0;
    const int y = valid ? (int)(row % N) : 0;
    const float* map = in + b * N * N;
    float2* s = lds + g * LdsSize<N>::value;
    auto gs = [&](int i, float2 v) { s[lds_pad(i)] = v; };
    // When a workgroup's G rows are consecutive rows of one map, those rows and their neighbours above
    // and below are first staged into the (still unused) LDS area with coalesced float4 loads: the
    // transform's first stage, which evaluates the right-hand side, reads every input before its first
    // write into the area, so the staged rows and the transform can share it.
    constexpr bool STAGE = N % G == 0 && (G + 2) * N <= 2 * G * LdsSize<N>::value && N % 4 == 0;
    if constexpr (STAGE) {
        float* st = reinterpret_cast<float*>(lds);
        const long long row0 = (long long)blockIdx.x * G;
        const bool wg_valid = row0 < (long long)batch * N;  // N % G == 0: all G rows valid or none
        const int y0 = (int)(row0 % N);
        const float* m0 = in + (wg_valid ? row0 / N : 0) * N * N;
        const int r_lo = y0 == 0 ? 1 : 0, r_hi = (y0 + G == N) ? G : G + 1;  // staged rows r_lo .. r_hi
        if (wg_valid) {
            const int nv = (r_hi - r_lo + 1) * (N / 4);
            const float4* src = reinterpret_cast<const float4*>(m0 + (long long)(y0 - 1 + r_lo) * N);
            float4* dst = reinterpret_cast<float4*>(st + r_lo * N);
            for (int e = threadIdx.x; e < nv; e += G * T) dst[e] = src[e];
        }
        __syncthreads();
        auto gl = [&](int n) -> float2 {
            if (!valid) return make_float2(0.f, 0.f);
            const int x = mk_src(n, N);
            return make_float2(mode == 0 ? rhs_staged(st, N, x, y, y0, sincos) : st[(y - y0 + 1) * N + x], 0.f);
        };
        Plan<N>::template run<-1>(t, s, tw_n, gl, gs);
    } else {
        auto gl = [&](int n) -> float2 {
            if (!valid) return make_float2(0.f, 0.f);
            const int x = mk_src(n, N);
            return make_float2(mode == 0 ? rhs_at(map, N, x, y, sincos) : map[x + N * y], 0.f);
        };
        Plan<N>::template run<-1>(t, s, tw_n, gl, gs);
    }
    __syncthreads();
    if (!valid) return;
    const float s0 = sqrtf(1.f / N), s1 = sqrtf(2.f / N);
    float* dst = out + row * N;
    for (int k = t; k < N; k += T) {
        const float2 V = s[lds_pad(k)];
        const float2 w = tw_4n[k];  // exp(-i pi k / 2N)
        dst[k] = (V.x * w.x - V.y * w.y) * (k == 0 ? s0 : s1);
    }
}

// DCT-III of one line held as real coefficients `X` (LDS, N floats) into `s`, final values handed to
// put(n, value). refcompat selects the reference's literal 2N construction (requires LdsSize<2N>).
template <int N, int DIR_UNUSED = 0>
struct Dct3 {
    template <class Put>
    __device__ __forceinline__ static void exact(int t, const float* X, float2* s, const float2* tw_n,
                                                 const float2* tw_4n, Put& put)
    {
        // 1 / sqrt(1/N) and 1 / sqrt(2/N) (compile-time constants): the un-normalisation multiplies
        const float r0 = 1.f / sqrtf(1.f / N), r1 = 1.f / sqrtf(2.f / N);
        auto gl = [&](int k) -> float2 {
            const float ck = X[k] * (k == 0 ? r0 : r1);
            const float cn = (k == 0) ? 0.f : X[N - k] * r1;
            const float2 w = tw_4n[k];  // exp(-i pi k/2N); we need exp(+i pi k/2N) = conj
            // (ck - i cn) * (w.x - i w.y)
            return make_float2(ck * w.x - cn * w.y, -ck * w.y - cn * w.x);
        };
        const float inv = 1.f / N;
        auto gs = [&](int n, float2 v) {
            const int x = mk_src(n, N);
            put(x, v.x * inv);
        };
        Plan<N>::template run<+1>(t, s, tw_n, gl, gs);
    }
};

// Reference IDCT of one line (countIDCT + inverse 2N FFT + cufft2float), dTemp = 1/sqrt(2N).
template <int N, class Put>
__device__ __forceinline__ void idct_refcompat(int t, const float* X, float2* s, const float2* tw_2n,
                                               const float2* tw_4n, Put& put)
{
    const float dT = 1.f / sqrtf(2.f * N);
    const float r2 = 1.41421356237309515f;
    auto gl = [&](int i) -> float2 {
        if (i < N) {
            const float2 w = tw_4n[i];  // (cos, -sin)(pi i / 2N)
            const float C = w.x, S = -w.y, V = X[i] * dT;
            return i == 0 ? make_float2(C * V * r2, S * V * r2) : make_float2(C * V, S * V);
        }
        const int ii = i - N;
        if (ii == 0) return make_float2(0.f, 0.f);
        const float2 w = tw_4n[ii + 1];  // angle (ii+1) pi / 2N
        const float S2 = -w.y, C2 = -w.x, V2 = X[N - ii] * dT;
        return make_float2(S2 * V2, C2 * V2);
    };
    auto gs = [&](int i, float2 v) {
        if (i < N) put(i, v.x);
    };
    Plan<2 * N>::template run<+1>(t, s, tw_2n, gl, gs);
}

// ---------------------------------------------------------------- 2. columns: DCT-II, / lambda, DCT-III
// LDS per column: one complex area (LdsSize<LN>). The N real coefficients between the two transforms
// are held in registers while the area is read, then written into it as floats: the DCT-III reads them
// only in its first stage, before its first write to the area. One area per column keeps 60 maps of
// 256 columns within one round of workgroups.
template <int N, int C, int LN>
__global__ __launch_bounds__(C* Plan<LN>::T) void k_uw_cols(const float* __restrict__ in, int batch,
                                                            unsigned flags, const double* __restrict__ eig,
                                                            const float2* __restrict__ tw_n,
                                                            const float2* __restrict__ tw_2n,
                                                            const float2* __restrict__ tw_4n,
                                                            float* __restrict__ out)
{
    constexpr int TL = Plan<LN>::T;  // threads per column (sized for the longest transform used)
    constexpr int LS = LdsStride<LN, C>::value;
    static_assert(2 * LdsSize<LN>::value >= N, "coefficients must fit the complex area");
    constexpr int J = (N + TL - 1) / TL;  // coefficients per thread
    __shared__ float2 lds[C * LS];
    const int cidx = threadIdx.x % C, t = threadIdx.x / C;
    const int tiles = N / C;
    const long long b = blockIdx.x / tiles;
    const int kx = (blockIdx.x % tiles) * C + cidx;
    const float* col = in + b * N * N + kx;
    float2* s = lds + cidx * LS;
    float* X = reinterpret_cast<float*>(s);  // valid once the forward transform's output has been read
    auto gl = [&](int n) -> float2 { return make_float2(col[(long long)mk_src(n, N) * N], 0.f); };
    auto gs = [&](int i, float2 v) { s[lds_pad(i)] = v; };
    Plan<N>::template run<-1>(t, s, tw_n, gl, gs);
    __syncthreads();
    const float s0 = sqrtf(1.f / N), s1 = sqrtf(2.f / N);
    const bool refeig = flags & F_REF_EIG;
    float xv[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int ky = t + j * TL;
        if (t < TL && ky < N) {
            const float2 V = s[lds_pad(ky)];
            const float2 w = tw_4n[ky];
            const float Xv = (V.x * w.x - V.y * w.y) * (ky == 0 ? s0 : s1);
            double lam;
            if (refeig)
                lam = eig[kx] + eig[ky] - 4.0;
            else
                lam = -4.0 * (eig[kx] + eig[ky]);
            // Xv / lambda as a product with the hardware reciprocal (v_rcp_f32, about 1 ulp) of the
            // eigenvalue sum formed in double: two ulps at most, in place of a double division
            xv[j] = (kx == 0 && ky == 0) ? 0.f : Xv * __builtin_amdgcn_rcpf((float)lam);
        }
    }
    __syncthreads();  // every read of the transform's output precedes the coefficients' writes
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int ky = t + j * TL;
        if (t < TL && ky < N) X[ky] = xv[j];
    }
    __syncthreads();
    float* dcol = out + b * N * N + kx;
    auto put = [&](int y, float v) { dcol[(long long)y * N] = v; };
    if (flags & F_REF_IDCT) {
        if constexpr (LN == 2 * N) idct_refcompat<N>(t, X, s, tw_2n, tw_4n, put);
    } else {
        Dct3<N>::exact(t, X, s, tw_n, tw_4n, put);
    }
}

// ---------------------------------------------------------------- 3. rows: DCT-III along x
template <int N, int G, int LN>
__global__ __launch_bounds__(G* Plan<LN>::T) void k_uw_rows_inv(const float* __restrict__ in, int batch,
                                                                unsigned flags,
                                                                const float* __restrict__ wrapped,
                                                                const float2* __restrict__ tw_n,
                                                                const float2* __restrict__ tw_2n,
                                                                const float2* __restrict__ tw_4n,
                                                                float* __restrict__ out)
{
    constexpr int TL = Plan<LN>::T;
    __shared__ float2 lds[G * LdsSize<LN>::value];
    const int g = threadIdx.x / TL, t = threadIdx.x % TL;
    const long long row = (long long)blockIdx.x * G + g;
    const bool valid = row < (long long)batch * N;
    const long long r = valid ? row : 0;
    float2* s = lds + g * LdsSize<LN>::value;
    // The DCT-III's first stage reads coefficients k and N - k of the row straight from the input (the
    // row's 1 KiB stays in cache); no LDS copy, so 60 maps of 256 rows fit one round of workgroups.
    const float* X = in + r * N;
    const bool cong = (flags & F_CONGRUENT) && wrapped;
    auto put = [&](int x, float v) {
        if (!valid) return;
        if (cong) {
            const float w = wrapped[r * N + x];
            v = w + 2.f * kPI * rintf((v - w) / (2.f * kPI));
        }
        out[r * N + x] = v;
    };
    if (flags & F_REF_IDCT) {
        if constexpr (LN == 2 * N) idct_refcompat<N>(t, X, s, tw_2n, tw_4n, put);
    } else {
        Dct3<N>::exact(t, X, s, tw_n, tw_4n, put);
    }
}

// Rows of the last unwrapping pass fused with the row pass of the Rytov spectrum (k_rytov_rows of the
// reconstruction module, same transform and twiddles): each row's DCT-III result, after the congruence
// step, goes to the unwrapped map (when wanted) and stays in the row's LDS area as floats (the DCT-III's
// last stage writes only after its barrier), where the forward row transform of u = logamp + i phi
// reads it. Row spectra: P[r][k].
template <int N, int G>
__global__ __launch_bounds__(G* Plan<N>::T) void k_uw_rows_inv_rytov(const float* __restrict__ in, int batch,
                                                                      unsigned flags,
                                                                      const float* __restrict__ wrapped,
                                                                      const float* __restrict__ la,
                                                                      const float2* __restrict__ tw_n,
                                                                      const float2* __restrict__ tw_4n,
                                                                      float* __restrict__ out,
                                                                      float2* __restrict__ P)
{
    constexpr int TL = Plan<N>::T;
    static_assert(2 * LdsSize<N>::value >= N, "the phase row must fit the complex area");
    __shared__ float2 lds[G * LdsSize<N>::value];
    const int g = threadIdx.x / TL, t = threadIdx.x % TL;
    const long long row = (long long)blockIdx.x * G + g;
    const bool valid = row < (long long)batch * N;
    const long long r = valid ? row : 0;
    float2* s = lds + g * LdsSize<N>::value;
    float* ph = reinterpret_cast<float*>(s);
    const float* X = in + r * N;
    const bool cong = (flags & F_CONGRUENT) && wrapped;
    auto put = [&](int x, float v) {
        if (cong) {
            const float w = wrapped[r * N + x];
            v = w + 2.f * kPI * rintf((v - w) / (2.f * kPI));
        }
        ph[x] = v;
        if (valid && out) out[r * N + x] = v;
    };
    Dct3<N>::exact(t, X, s, tw_n, tw_4n, put);
    __syncthreads();
    auto gl = [&](int x) -> float2 { return make_float2(la[r * N + x], ph[x]); };
    auto gs = [&](int k, float2 v) {
        if (valid) P[r * N + k] = v;
    };
    Plan<N>::template run<-1>(t, s, tw_n, gl, gs);
}

// Adjoint helper: (weighted) zero-flux Laplacian, g_in = sum_nbr w_pn (v_n - v_p); weights
// cos(phi_n - phi_p) for SINCOS, 1 otherwise.
__global__ void k_uw_stencil(const float* __restrict__ v, const float* __restrict__ phi, int batch,
                             int N, int sincos, float* __restrict__ out)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)batch * N * N) return;
    const long long b = g / ((long long)N * N);
    const int k = (int)(g % ((long long)N * N)), i = k % N, j = k / N;
    const float* vv = v + b * N * N;
    const float* pp = phi + b * N * N;
    float acc = 0.f;
    auto term = [&](int kn) {
        const float w = sincos ? cosf(pp[kn] - pp[k]) : 1.f;
        acc += w * (vv[kn] - vv[k]);
    };
    if (i < N - 1) term(k + 1);
    if (i > 0) term(k - 1);
    if (j < N - 1) term(k + N);
    if (j > 0) term(k - N);
    out[g] = acc;
}

// ---------------------------------------------------------------- FastUnwrapping's map conventions
// One workgroup per map: minimum (and maximum) of its n values, left in mn[0] (mx[0]) for every thread.
template <int T>
__device__ __forceinline__ void map_min_max(const float* __restrict__ m, int n, float* red, float& mn, float& mx)
{
    float lo = INFINITY, hi = -INFINITY;
    for (int i = threadIdx.x; i < n; i += T) {
        const float v = m[i];
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    red[threadIdx.x] = lo;
    red[T + threadIdx.x] = hi;
    __syncthreads();
    for (int w = T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[threadIdx.x] = fminf(red[threadIdx.x], red[threadIdx.x + w]);
            red[T + threadIdx.x] = fmaxf(red[T + threadIdx.x], red[T + threadIdx.x + w]);
        }
        __syncthreads();
    }
    mn = red[0];
    mx = red[T];
}

// PhaseNormalize + RescalePhaseMap (FastUnwrapping.cu:119-142): each map rescaled to
// 2 pi (x - min) / (max - min), evaluated in double as the reference's double-precision M_PI promotes it.
__global__ __launch_bounds__(256) void k_uw_normalize(const float* __restrict__ in, float* __restrict__ out, int n)
{
    __shared__ float red[512];
    const float* m = in + (size_t)blockIdx.x * n;
    float* o = out + (size_t)blockIdx.x * n;
    float mn, mx;
    map_min_max<256>(m, n, red, mn, mx);
    const double span = (double)(mx - mn);
    for (int i = threadIdx.x; i < n; i += 256) o[i] = (float)(2.0 * M_PI * (double)(m[i] - mn) / span);
}

// subtractMin + PhaseMapSubMin (FastUnwrapping.cu:346-369): each map minus its own minimum, in place.
__global__ __launch_bounds__(256) void k_uw_sub_min(float* __restrict__ x, int n)
{
    __shared__ float red[512];
    float* m = x + (size_t)blockIdx.x * n;
    float mn, mx;
    map_min_max<256>(m, n, red, mn, mx);
    __syncthreads();  // every thread has read the minimum before any value changes
    for (int i = threadIdx.x; i < n; i += 256) m[i] -= mn;
}

// ---------------------------------------------------------------- host side
static const double* eig_table(Ctx& c, int N, bool refcompat, bool existing = false)
{
    const std::string name = std::string(refcompat ? "uw.eigref." : "uw.eig.") + std::to_string(N);
    auto it = c.scratch.find(name);
    if (it != c.scratch.end() && it->second.ptr) return (const double*)it->second.ptr;
    CTDM_REQUIRE(!existing, "unwrap: eigenvalue table missing (it is built when the context is created)");
    std::vector<double> h(N);
    for (int k = 0; k < N; ++k) {
        if (refcompat)
            h[k] = 2.0 * std::cos(4.0 * M_PI * k / (4.0 * N - 1));
        else {
            const double sv = std::sin(M_PI * k / (2.0 * N));
            h[k] = sv * sv;
        }
    }
    double* d = (double*)c.buf(name, sizeof(double) * N, true);  // one table for both lanes
    CTDM_HIP(hipMemcpy(d, h.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    return d;
}

template <int N, int LN>
static void run_cols(Ctx& c, const float* in, int batch, unsigned flags, float* out, bool existing)
{
    constexpr int C = (256 / Plan<LN>::T) < N ? (256 / Plan<LN>::T) : N;
    const double* eig = eig_table(c, N, flags & F_REF_EIG, existing);
    const float2* tw2 = (LN == 2 * N) ? c.tw(2 * N) : c.tw(N);
    hipLaunchKernelGGL((k_uw_cols<N, C, LN>), dim3(batch * (N / C)), dim3(C * Plan<LN>::T), 0, c.stream,
                       in, batch, flags, eig, c.tw(N), tw2, c.tw(4 * N), out);
    CTDM_CHECK_LAUNCH("k_uw_cols");
}

template <int N, int LN>
static void run_rows_inv(Ctx& c, const float* in, int batch, unsigned flags, const float* wrapped,
                         float* out)
{
    constexpr int G = 256 / Plan<LN>::T;
    const float2* tw2 = (LN == 2 * N) ? c.tw(2 * N) : c.tw(N);
    hipLaunchKernelGGL((k_uw_rows_inv<N, G, LN>), dim3(grid1d((long long)batch * N, G)), dim3(G * Plan<LN>::T),
                       0, c.stream, in, batch, flags, wrapped, c.tw(N), tw2, c.tw(4 * N), out);
    CTDM_CHECK_LAUNCH("k_uw_rows_inv");
}

template <int N>
static void run_rows_fwd(Ctx& c, const float* in, int batch, int mode, bool sincos, float* out)
{
    constexpr int G = 256 / Plan<N>::T;
    hipLaunchKernelGGL((k_uw_rows_fwd<N, G>), dim3(grid1d((long long)batch * N, G)), dim3(G * Plan<N>::T), 0,
                       c.stream, in, batch, mode, (int)sincos, c.tw(N), c.tw(4 * N), out);
    CTDM_CHECK_LAUNCH("k_uw_rows_fwd");
}

// Intermediate spectra of one solve: context scratch, or (allocation-free entry point) the fixed
// device array. `existing` also forbids building a missing table.
struct Work {
    float* D1 = nullptr;
    float* D2 = nullptr;
    bool existing = false;
};

// Whole solve for one size: mode 0 (wrapped input) or 1 (raw input, adjoint).
template <int N>
static void solve(Ctx& c, const float* in, int batch, unsigned flags, int mode, const float* wrapped,
                  float* out, Work w)
{
    if (!w.D1) {
        w.D1 = (float*)c.buf("uw.D1", sizeof(float) * (size_t)batch * N * N);
        w.D2 = (float*)c.buf("uw.D2", sizeof(float) * (size_t)batch * N * N);
    }
    run_rows_fwd<N>(c, in, batch, mode, flags & F_SINCOS, w.D1);
    if (flags & F_REF_IDCT) {
        if constexpr (N <= 256) {
            run_cols<N, 2 * N>(c, w.D1, batch, flags, w.D2, w.existing);
            run_rows_inv<N, 2 * N>(c, w.D2, batch, flags, wrapped, out);
        } else {
            fail(CTDM_EUNSUPPORTED, "REFCOMPAT_IDCT needs N <= 256");
        }
    } else {
        run_cols<N, N>(c, w.D1, batch, flags, w.D2, w.existing);
        run_rows_inv<N, N>(c, w.D2, batch, flags, wrapped, out);
    }
}

static void solve_any(Ctx& c, const float* in, int batch, int N, unsigned flags, int mode,
                      const float* wrapped, float* out, Work w = Work{})
{
    switch (N) {
    case 32: solve<32>(c, in, batch, flags, mode, wrapped, out, w); break;
    case 64: solve<64>(c, in, batch, flags, mode, wrapped, out, w); break;
    case 128: solve<128>(c, in, batch, flags, mode, wrapped, out, w); break;
    case 192: solve<192>(c, in, batch, flags, mode, wrapped, out, w); break;
    case 256: solve<256>(c, in, batch, flags, mode, wrapped, out, w); break;
    default: fail(CTDM_EUNSUPPORTED, "unwrap size " + std::to_string(N) + " has no kernel");
    }
}

// ---------------------------------------------------------------- allocation-free entry point
// The two intermediate spectra of ctdm_unwrap_dct_noalloc, CTDM_NOALLOC_MAX_ELEMS floats each, in one
// file-scope device array. Every device the library runs on holds its own copy until the process ends.
// In host code `scratch` is a host-side stand-in: kernels get the per-device address that
// hipGetSymbolAddress returns (noalloc_setup), never the array's name.
static __device__ __align__(256) float scratch[2 * (size_t)CTDM_NOALLOC_MAX_ELEMS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_setup(Ctx& c)
{
    CTDM_REQUIRE(c.device >= 0 && c.device < kMaxDevices, "device ordinal out of range");
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[c.device]) {  // once per device, with that device current
            float* ptr = nullptr;
            CTDM_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
            CTDM_REQUIRE(ptr != nullptr, "scratch symbol has no device address");
            g_scratch_ptr[c.device] = ptr;
        }
        c.noalloc_scratch = g_scratch_ptr[c.device];
    }
    // Every table the unwrap kernels read, built now so that the allocation-free call finds them.
    for (int N : {32, 64, 128, 192, 256}) {
        eig_table(c, N, false);
        eig_table(c, N, true);
        c.tw(N);
        c.tw(2 * N);
        c.tw(4 * N);
    }
}

void unwrap_dct_noalloc(Ctx& c, const float* wrapped, float* unwrapped, int batch, int M, unsigned flags)
{
    CTDM_REQUIRE(wrapped && unwrapped, "unwrap_dct_noalloc: null pointer");
    CTDM_REQUIRE(batch >= 0 && M > 0, "unwrap_dct_noalloc: bad size");
    CTDM_REQUIRE(wrapped != unwrapped, "unwrap_dct_noalloc: in-place call not supported");
    CTDM_REQUIRE((flags & ~0xfu) == 0, "unwrap_dct_noalloc: unknown flag");
    CTDM_REQUIRE((size_t)batch * M * M <= (size_t)CTDM_NOALLOC_MAX_ELEMS,
                 "unwrap_dct_noalloc: batch * M * M = " + std::to_string((size_t)batch * M * M) +
                     " exceeds the fixed scratch of " + std::to_string((size_t)CTDM_NOALLOC_MAX_ELEMS) +
                     " elements");
    CTDM_REQUIRE(c.noalloc_scratch, "unwrap_dct_noalloc: scratch address not set up");
    if (batch == 0) return;
    Work w;
    w.D1 = c.noalloc_scratch;
    w.D2 = c.noalloc_scratch + (size_t)CTDM_NOALLOC_MAX_ELEMS;
    w.existing = true;
    solve_any(c, wrapped, batch, M, flags, 0, wrapped, unwrapped, w);
}

// With CTDM_UNWRAP_REFCOMPAT_FAST the maps are first rescaled to [0, 2 pi] (into scratch `name`) and
// the result of each map is shifted to a zero minimum, FastUnwrapping's conventions.
static const float* normalized_input(Ctx& c, const float* wrapped, int batch, int M, const char* name)
{
    float* norm = (float*)c.buf(name, sizeof(float) * (size_t)batch * M * M);
    hipLaunchKernelGGL(k_uw_normalize, dim3(batch), dim3(256), 0, c.stream, wrapped, norm, M * M);
    CTDM_CHECK_LAUNCH("k_uw_normalize");
    return norm;
}

static void subtract_min(Ctx& c, float* out, int batch, int M)
{
    hipLaunchKernelGGL(k_uw_sub_min, dim3(batch), dim3(256), 0, c.stream, out, M * M);
    CTDM_CHECK_LAUNCH("k_uw_sub_min");
}

void unwrap_dct_batched(Ctx& c, const float* wrapped, float* unwrapped, int batch, int M,
                        unsigned flags)
{
    CTDM_REQUIRE(wrapped && unwrapped, "unwrap_dct: null pointer");
    CTDM_REQUIRE(batch >= 0 && M > 0, "unwrap_dct: bad size");
    CTDM_REQUIRE(wrapped != unwrapped, "unwrap_dct: in-place call not supported");
    CTDM_REQUIRE((flags & ~F_ALL) == 0, "unwrap_dct: unknown flag");
    if (batch == 0) return;
    if (flags & F_REF_FAST) {
        const float* norm = normalized_input(c, wrapped, batch, M, "uw.norm");
        solve_any(c, norm, batch, M, flags & ~F_REF_FAST, 0, norm, unwrapped);
        subtract_min(c, unwrapped, batch, M);
        return;
    }
    solve_any(c, wrapped, batch, M, flags, 0, wrapped, unwrapped);
}

template <int N>
static void uw_rytov(Ctx& c, const float* wrapped, float* unwrapped, const float* logamp, int batch, unsigned flags,
                     float2* P)
{
    constexpr int G = 256 / Plan<N>::T;
    float* D1 = (float*)c.buf("uw.D1", sizeof(float) * (size_t)batch * N * N);
    float* D2 = (float*)c.buf("uw.D2", sizeof(float) * (size_t)batch * N * N);
    run_rows_fwd<N>(c, wrapped, batch, 0, flags & F_SINCOS, D1);
    run_cols<N, N>(c, D1, batch, flags, D2, false);
    hipLaunchKernelGGL((k_uw_rows_inv_rytov<N, G>), dim3(grid1d((long long)batch * N, G)), dim3(G * Plan<N>::T), 0,
                       c.stream, D2, batch, flags, wrapped, logamp, c.tw(N), c.tw(4 * N), unwrapped, P);
    CTDM_CHECK_LAUNCH("k_uw_rows_inv_rytov");
}

void unwrap_dct_rytov_rows(Ctx& c, const float* wrapped, float* unwrapped, const float* logamp, int batch, int M,
                           unsigned flags, float2* P)
{
    CTDM_REQUIRE(wrapped && logamp && P, "unwrap_dct_rytov_rows: null pointer");
    CTDM_REQUIRE(batch >= 0 && M > 0, "unwrap_dct_rytov_rows: bad size");
    CTDM_REQUIRE(wrapped != unwrapped, "unwrap_dct_rytov_rows: in-place call not supported");
    CTDM_REQUIRE((flags & ~F_ALL) == 0 && !(flags & (F_REF_FAST | F_REF_IDCT)),
                 "unwrap_dct_rytov_rows: unsupported flag");
    if (batch == 0) return;
    switch (M) {
    case 32: uw_rytov<32>(c, wrapped, unwrapped, logamp, batch, flags, P); break;
    case 64: uw_rytov<64>(c, wrapped, unwrapped, logamp, batch, flags, P); break;
    case 128: uw_rytov<128>(c, wrapped, unwrapped, logamp, batch, flags, P); break;
    case 192: uw_rytov<192>(c, wrapped, unwrapped, logamp, batch, flags, P); break;
    case 256: uw_rytov<256>(c, wrapped, unwrapped, logamp, batch, flags, P); break;
    default: fail(CTDM_EUNSUPPORTED, "unwrap size " + std::to_string(M) + " has no kernel");
    }
}

void unwrap_dct_adjoint(Ctx& c, const float* wrapped, const float* g_out, float* g_in, int batch,
                        int M, unsigned flags)
{
    CTDM_REQUIRE(wrapped && g_out && g_in, "unwrap_dct_adjoint: null pointer");
    CTDM_REQUIRE(!(flags & (F_REF_IDCT | F_REF_FAST)),
                 "unwrap_dct_adjoint: the reference IDCT and FastUnwrapping's rescaling have no adjoint here");
    if (batch == 0) return;
    const size_t n = (size_t)batch * M * M;
    if (flags & F_CONGRUENT) {  // output = wrapped + 2 pi k with k piecewise constant: J = I
        CTDM_HIP(hipMemcpyAsync(g_in, g_out, sizeof(float) * n, hipMemcpyDeviceToDevice, c.stream));
        return;
    }
    // J = D_rho * P with P the (symmetric) Poisson operator and D_rho the (symmetric) linearised
    // RHS stencil, so J^T g = D_rho (P g).
    float* tmp = (float*)c.buf("uw.adj", sizeof(float) * n);
    solve_any(c, g_out, batch, M, flags & F_REF_EIG, 1, nullptr, tmp);
    hipLaunchKernelGGL(k_uw_stencil, dim3(grid1d((long long)n, 256)), dim3(256), 0, c.stream, tmp, wrapped, batch,
                       M, (int)(flags & F_SINCOS), g_in);
    CTDM_CHECK_LAUNCH("k_uw_stencil");
}

// ---------------------------------------------------------------- plain (reference-structured) path
// NumericDerivative1 (DCT_Unwrapping.cu:360-392)
__global__ void k_numeric_derivative(float* __restrict__ ox, float* __restrict__ oy, const float* __restrict__ in,
                                     int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N || j >= N) return;
    const int k = i + N * j;
    const float tx = (i == N - 1) ? -in[k] : in[k + 1] - in[k];
    const float ty = (j == N - 1) ? -in[k] : in[k + N] - in[k];
    ox[k] = wrapdiff(tx);
    oy[k] = wrapdiff(ty);
}

// SumDerivative (DCT_Unwrapping.cu:394-417)
__global__ void k_sum_derivative(float* __restrict__ o, const float* __restrict__ ox, const float* __restrict__ oy,
                                 int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N || j >= N) return;
    const int k = i + N * j;
    const float c1 = (i == N - 1) ? 0.f : ox[k];
    const float c2 = (i == 0) ? 0.f : ox[k - 1];
    const float c3 = (j == N - 1) ? 0.f : oy[k];
    const float c4 = (j == 0) ? 0.f : oy[k - N];
    o[k] = (c1 - c2) + (c3 - c4);
}

// sin/cos RHS for the plain path: one stencil kernel.
__global__ void k_sincos_rhs(float* __restrict__ o, const float* __restrict__ in, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N || j >= N) return;
    o[i + N * j] = rhs_at(in, N, i, j, true);
}

// float2cufft (DCT_Unwrapping.cu:144-162): row mirror to length 2N.
__global__ void k_mirror(float2* __restrict__ o, const float* __restrict__ in, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N || j >= N) return;
    const float v = in[i + N * j];
    o[i + 2 * N * j] = make_float2(v, 0.f);
    o[(2 * N - 1 - i) + 2 * N * j] = make_float2(v, 0.f);
}

// countFDCT (DCT_Unwrapping.cu:202-216)
__global__ void k_count_fdct(float* __restrict__ o, const float2* __restrict__ in, int N, float dT)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N || j >= N) return;
    const float2 v = in[i + 2 * N * j];
    const float a = (float)i * kPI / (2.f * N);
    o[i + N * j] = (i == 0) ? v.x * dT / 1.41421356237309515f : (v.x * cosf(a) + v.y * sinf(a)) * dT;
}

// transposeShared (DCT_Unwrapping.cu:179-200), square N x N with 32 x 32 LDS tiles.
__global__ void k_transpose(float* __restrict__ o, const float* __restrict__ in, int N)
{
    __shared__ float tile[32][33];
    const int x = blockIdx.x * 32 + threadIdx.x;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int y = blockIdx.y * 32 + r;
        if (x < N && y < N) tile[r][threadIdx.x] = in[x + (long long)N * y];
    }
    __syncthreads();
    const int x2 = blockIdx.y * 32 + threadIdx.x;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int y2 = blockIdx.x * 32 + r;
        if (x2 < N && y2 < N) o[x2 + (long long)N * y2] = tile[threadIdx.x][r];
    }
}

// devConstant (DCT_Unwrapping.cu:300-312) with the eigenvalue choice of the batched path.
__global__ void k_dev_constant(float* __restrict__ buf, int N, int refeig, const double* __restrict__ eig)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N || j >= N) return;
    const int k = i + N * j;
    if (i == 0 && j == 0) {
        buf[k] = 0.f;
        return;
    }
    const double lam = refeig ? eig[i] + eig[j] - 4.0 : -4.0 * (eig[i] + eig[j]);
    buf[k] = (float)((double)buf[k] / lam);
}

// countIDCT (DCT_Unwrapping.cu:218-246); `shift` = 1 reproduces the reference's (i+1) twiddle.
__global__ void k_count_idct(float2* __restrict__ o, const float* __restrict__ in, int N, float dT, int shift)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N || j >= N) return;
    const float a = (float)i * kPI / (2.f * N);
    const float C = cosf(a), S = sinf(a), V = in[i + N * j] * dT;
    const float a2 = ((float)i + (float)shift) * kPI / (2.f * N);
    const float S2 = sinf(a2), C2 = -cosf(a2);
    const float r2 = 1.41421356237309515f;
    float2 t1 = (i == 0) ? make_float2(C * V * r2, S * V * r2) : make_float2(C * V, S * V);
    const float V2 = (i == 0) ? 0.f : in[(N - i) + N * j] * dT;
    float2 t2 = (i == 0) ? make_float2(0.f, 0.f) : make_float2(S2 * V2, C2 * V2);
    o[i + 2 * N * j] = t1;
    o[i + 2 * N * j + N] = t2;
}

// cufft2float (DCT_Unwrapping.cu:164-177)
__global__ void k_real_part(float* __restrict__ o, const float2* __restrict__ in, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N || j >= N) return;
    o[i + N * j] = in[i + 2 * N * j].x;
}

__global__ void k_congruent(float* __restrict__ o, const float* __restrict__ w, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = w[i] + 2.f * kPI * rintf((o[i] - w[i]) / (2.f * kPI));
}

void unwrap_dct_plain_one(Ctx& c, const float* wrapped, float* out, int N, unsigned flags)
{
    CTDM_REQUIRE(N == 32 || N == 64 || N == 128 || N == 192 || N == 256, "plain unwrap: N in {32, 64, 128, 192, 256}");
    if (flags & F_REF_FAST) {
        const float* norm = normalized_input(c, wrapped, 1, N, "puw.norm");
        unwrap_dct_plain_one(c, norm, out, N, flags & ~F_REF_FAST);
        subtract_min(c, out, 1, N);
        return;
    }
    const size_t n = (size_t)N * N;
    float* inX = (float*)c.buf("puw.inX", sizeof(float) * n);
    float* inY = (float*)c.buf("puw.inY", sizeof(float) * n);
    float* lap = (float*)c.buf("puw.lap", sizeof(float) * n);
    float* dst = (float*)c.buf("puw.dst", sizeof(float) * n);
    float2* cplx = (float2*)c.buf("puw.cplx", sizeof(float2) * 2 * n);
    const double* eig = eig_table(c, N, flags & F_REF_EIG);
    dim3 blk(32, 8), grd((N + 31) / 32, (N + 7) / 8), tg(N / 32, N / 32);
    const float dT = 1.f / sqrtf(2.f * N);
    if (flags & F_SINCOS) {
        hipLaunchKernelGGL(k_sincos_rhs, grd, blk, 0, c.stream, lap, wrapped, N);
    } else {
        hipLaunchKernelGGL(k_numeric_derivative, grd, blk, 0, c.stream, inX, inY, wrapped, N);
        hipLaunchKernelGGL(k_sum_derivative, grd, blk, 0, c.stream, lap, inX, inY, N);
    }
    // myDCT (DCT_Unwrapping.cu:61-83): two passes of mirror / 2N FFT / countFDCT / transpose
    for (int pass = 0; pass < 2; ++pass) {
        float* src = pass == 0 ? lap : inX;
        hipLaunchKernelGGL(k_mirror, grd, blk, 0, c.stream, cplx, src, N);
        fft_rows(c, cplx, 2 * N, N, -1, 1.f, false);
        hipLaunchKernelGGL(k_count_fdct, grd, blk, 0, c.stream, dst, cplx, N, dT);
        hipLaunchKernelGGL(k_transpose, tg, dim3(32, 8), 0, c.stream, inX, dst, N);
    }
    // devConstant
    hipLaunchKernelGGL(k_dev_constant, grd, blk, 0, c.stream, inX, N, (int)((flags & F_REF_EIG) != 0), eig);
    // myIDCT (DCT_Unwrapping.cu:85-107): two passes of transpose / countIDCT / inverse 2N FFT / Re
    for (int pass = 0; pass < 2; ++pass) {
        float* src = pass == 0 ? inX : inY;
        hipLaunchKernelGGL(k_transpose, tg, dim3(32, 8), 0, c.stream, dst, src, N);
        hipLaunchKernelGGL(k_count_idct, grd, blk, 0, c.stream, cplx, dst, N, dT, (int)((flags & F_REF_IDCT) != 0));
        fft_rows(c, cplx, 2 * N, N, +1, 1.f, false);
        hipLaunchKernelGGL(k_real_part, grd, blk, 0, c.stream, pass == 0 ? inY : out, cplx, N);
    }
    if (flags & F_CONGRUENT)
        hipLaunchKernelGGL(k_congruent, dim3(grid1d((long long)n, 256)), dim3(256), 0, c.stream, out, wrapped, (int)n);
    CTDM_CHECK_LAUNCH("plain unwrap");
}

}  // namespace ctdm
