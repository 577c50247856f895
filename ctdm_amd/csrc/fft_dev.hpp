// Device-side FFT building blocks for gfx950: register DFT butterflies and an LDS-staged Stockham
// auto-sort transform executed cooperatively by T threads per transform.
//
// Convention: DIR = -1 is the forward transform exp(-2*pi*i*j*k/N) (cuFFT CUFFT_FORWARD), DIR = +1
// the unnormalised inverse exp(+2*pi*i*j*k/N) (CUFFT_INVERSE). Scaling is applied by the callers'
// epilogues (the reference scales in separate kernels, FFT.cu:853-916).
//
// Stockham stage (radix R, L = product of the radices already applied, S = N/R slots):
//   slot j, k = j mod L:  a_r = x[j + r*S] * W_{L*R}^{DIR*r*k},  a = DFT_R(a),
//                         y[(j - k)*R + k + r*L] = a_r
// Applying the stages for radices R0, R1, ... in order yields the DFT in natural order.
#pragma once
#include <hip/hip_runtime.h>

namespace ctdm {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)  // a * conj(b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Multiply by DIR*i: DIR = -1 -> -i, DIR = +1 -> +i.
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a)
{
    return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// cos / sin of 2*pi*k/32, k = 0..31 (exact to double rounding).
struct Tw32 {
    static constexpr double c[32] = {
        1.0, 0.98078528040323043, 0.92387953251128674, 0.83146961230254524, 0.70710678118654757,
        0.55557023301960218, 0.38268343236508984, 0.19509032201612833, 0.0, -0.19509032201612819,
        -0.38268343236508973, -0.55557023301960196, -0.70710678118654746, -0.83146961230254524,
        -0.92387953251128674, -0.98078528040323043, -1.0, -0.98078528040323043, -0.92387953251128674,
        -0.83146961230254546, -0.70710678118654768, -0.55557023301960218, -0.38268343236509034,
        -0.19509032201612866, 0.0, 0.19509032201612803, 0.38268343236509000, 0.55557023301960184,
        0.70710678118654735, 0.83146961230254524, 0.92387953251128652, 0.98078528040323043};
};

// In-register DFT of length R (power of two up to 32, 3 or 12), direction DIR.
template <int R, int DIR>
struct DFT;

template <int DIR>
struct DFT<1, DIR> {
    __device__ __forceinline__ static void run(float2*) {}
};

template <int DIR>
struct DFT<2, DIR> {
    __device__ __forceinline__ static void run(float2* a)
    {
        const float2 t = a[1];
        a[1] = csub(a[0], t);
        a[0] = cadd(a[0], t);
    }
};

template <int DIR>
struct DFT<3, DIR> {
    __device__ __forceinline__ static void run(float2* a)
    {
        // W = exp(DIR*2*pi*i/3) = -1/2 + DIR*i*sqrt(3)/2
        const float h = 0.86602540378443865f;
        const float2 t = cadd(a[1], a[2]);
        const float2 d = csub(a[1], a[2]);
        const float2 m = make_float2(a[0].x - 0.5f * t.x, a[0].y - 0.5f * t.y);
        const float2 r = mul_i<DIR>(cscale(d, h));
        a[0] = cadd(a[0], t);
        a[1] = cadd(m, r);
        a[2] = csub(m, r);
    }
};

template <int DIR>
struct DFT<4, DIR> {
    __device__ __forceinline__ static void run(float2* a)
    {
        const float2 t0 = cadd(a[0], a[2]);
        const float2 t1 = csub(a[0], a[2]);
        const float2 t2 = cadd(a[1], a[3]);
        const float2 t3 = mul_i<DIR>(csub(a[1], a[3]));
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    }
};

// Length 12 by the Good-Thomas prime-factor map (12 = 3 x 4, coprime, no twiddles): with
// n = (4 n1 + 3 n2) mod 12 and k = (4 k1 + 9 k2) mod 12, exp(2 pi i n k / 12) = W3^(n1 k1) W4^(n2 k2),
// so four-point DFTs over n2 followed by three-point DFTs over n1 give the twelve-point DFT.
template <int DIR>
struct DFT<12, DIR> {
    __device__ __forceinline__ static void run(float2* a)
    {
        float2 b[3][4];
#pragma unroll
        for (int n1 = 0; n1 < 3; ++n1) {
#pragma unroll
            for (int n2 = 0; n2 < 4; ++n2) b[n1][n2] = a[(4 * n1 + 3 * n2) % 12];
            DFT<4, DIR>::run(b[n1]);
        }
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            float2 c[3] = {b[0][k2], b[1][k2], b[2][k2]};
            DFT<3, DIR>::run(c);
#pragma unroll
            for (int k1 = 0; k1 < 3; ++k1) a[(4 * k1 + 9 * k2) % 12] = c[k1];
        }
    }
};

// Radix-2 decimation in time over two half-length register DFTs (R = 8, 16, 32).
template <int R, int DIR>
struct DFT {
    static_assert(R >= 8 && R <= 32 && (R & (R - 1)) == 0, "register DFT supports 2^k <= 32, 3 and 12");
    __device__ __forceinline__ static void run(float2* a)
    {
        float2 e[R / 2], o[R / 2];
#pragma unroll
        for (int k = 0; k < R / 2; ++k) {
            e[k] = a[2 * k];
            o[k] = a[2 * k + 1];
        }
        DFT<R / 2, DIR>::run(e);
        DFT<R / 2, DIR>::run(o);
#pragma unroll
        for (int k = 0; k < R / 2; ++k) {
            float2 t;
            if (k == 0) {
                t = o[k];
            } else if (4 * k == R) {
                t = mul_i<DIR>(o[k]);
            } else {
                const int m = k * (32 / R);
                const float c = (float)Tw32::c[m];
                const float s = (float)(DIR * Tw32::c[(m + 24) & 31]);  // sin(x) = cos(x - pi/2)
                t = cmul(o[k], make_float2(c, s));
            }
            a[k] = cadd(e[k], t);
            a[k + R / 2] = csub(e[k], t);
        }
    }
};

// LDS index with one float2 of padding every 16 elements (spreads stride-16 accesses over banks).
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
template <int N>
struct LdsSize {
    static constexpr int value = N + N / 16;
};

// Spacing (float2) of consecutive transforms' LDS areas. When the lanes of a wave walk the transform
// index fastest (LANES transforms side by side in each 16-lane group: the pencil and column kernels
// whose global accesses are coalesced across transforms), a spacing that is a multiple of 16 float2
// (128 B, every LDS bank once) would put all of them on one bank. Offsetting the spacing by 16 / LANES
// float2 makes (transform, slot) -> bank pair one-to-one for the padded radix-16 slot patterns
// 17 j + r and j + 17 r. LANES = 1 (the transform's own threads fastest) keeps the plain size.
template <int N, int LANES>
struct LdsStride {
    static constexpr int value =
        LANES <= 1 ? LdsSize<N>::value : (LdsSize<N>::value + 15) / 16 * 16 + (LANES >= 16 ? 1 : 16 / LANES);
};

// Copies a per-N twiddle table into LDS at kernel entry, padded like the data areas (dst holds
// LdsSize<N>::value float2). The copy overlaps the first data loads; the first Stockham stage uses
// no twiddles and its barrier publishes the table.
template <int N>
__device__ __forceinline__ void stage_twiddles(float2* dst, const float2* __restrict__ src)
{
    for (int i = threadIdx.x; i < N; i += blockDim.x) dst[lds_pad(i)] = src[i];
}

// View of an LDS twiddle table written by stage_twiddles. The padding keeps the lanes of one
// transform (slot k = lane) on distinct banks for the strided lookups m = r * k of the later stages.
struct TwPad {
    const float2* p;
    __device__ __forceinline__ float2 operator[](int m) const { return p[lds_pad(m)]; }
};

// Twiddle from a per-N table tw[m] = exp(-2*pi*i*m/N) (float2, computed in double on the host): a
// global pointer or a TwPad view.
template <int DIR, class TW>
__device__ __forceinline__ float2 twiddle(const TW& tw, int m)
{
    const float2 w = tw[m];
    return DIR > 0 ? make_float2(w.x, -w.y) : w;
}

// Input of a transform already in registers: the first stage of a radix-R0 transform of length N by T
// threads reads element j + r N/R0 (j = t + p T < N/R0) into its slot (p, r). A caller that loads those
// elements itself (to issue the loads ahead of other work) passes them as a RegFile in place of the load
// functor.
template <int P_, int R_>
struct RegFile {
    static constexpr int P = P_, R = R_;
    float2 v[P][R];
};
template <class X>
struct IsRegFile {
    static constexpr bool value = false;
};
template <int P, int R>
struct IsRegFile<RegFile<P, R>> {
    static constexpr bool value = true;
};
template <int N, int T, int R0>
using StageRegs = RegFile<(N / R0 + T - 1) / T, R0>;

// Complex value of one element of a transform's input: a complex sample, or a pair of consecutive
// real pixels (8- or 16-bit integers widened exactly) read as x[2n] + i x[2n+1].
__device__ __forceinline__ float2 as_complex(float2 v) { return v; }
__device__ __forceinline__ float2 as_complex(uchar2 v) { return make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ float2 as_complex(ushort2 v) { return make_float2((float)v.x, (float)v.y); }

// Loads, for thread t of a transform, the elements its first stage reads (element j + r N/R0 at
// p[(j + r N/R0) * stride], j = t + q T) into a RegFile.
template <int N, int T, int R0, class F, class E>
__device__ __forceinline__ void load_first_stage(F& f, const E* __restrict__ p, long long stride, int t)
{
    constexpr int S0 = N / R0;
#pragma unroll
    for (int q = 0; q < F::P; ++q) {
        const int j = t + q * T;
        if (S0 % T == 0 || j < S0) {
#pragma unroll
            for (int r = 0; r < F::R; ++r) f.v[q][r] = as_complex(p[(long long)(j + r * S0) * stride]);
        }
    }
}

// A per-N twiddle table copied into LDS in two steps by a workgroup of NT threads: fetch() loads this
// thread's entries into registers, put() writes them to LDS (padded like stage_twiddles). A kernel that
// issues its data loads between the two keeps them from queuing behind the wait for the table.
template <int N, int NT>
struct TwStager {
    static constexpr int K = (N + NT - 1) / NT;
    float2 r[K];
    __device__ __forceinline__ void fetch(const float2* __restrict__ src)
    {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = (int)threadIdx.x + k * NT;
            if (N % NT == 0 || i < N) r[k] = src[i];
        }
    }
    __device__ __forceinline__ void put(float2* dst) const
    {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = (int)threadIdx.x + k * NT;
            if (N % NT == 0 || i < N) dst[lds_pad(i)] = r[k];
        }
    }
};

// Cooperative N-point transform by T threads of one transform; threads with t >= T only take part
// in the barriers (so a workgroup sized for a longer transform can run a shorter one).
// gload(i) supplies input element i (called in the first stage), gstore(i, v) receives output
// element i (called in the last stage); intermediate stages exchange through the LDS area `s`
// (LdsSize<N>::value float2). Every thread of the workgroup must call this with the same template
// arguments: it contains __syncthreads().
template <int N, int T, int DIR, int L, int R, int... Rest>
struct Stockham {
    template <class TW, class GLoad, class GStore>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GLoad& gload, GStore& gstore)
    {
        constexpr int S = N / R;
        constexpr int P = (S + T - 1) / T;
        constexpr bool FIRST = (L == 1);
        constexpr bool LAST = (sizeof...(Rest) == 0);
        constexpr bool REG_IN = FIRST && IsRegFile<GLoad>::value;
        if constexpr (REG_IN) static_assert(GLoad::P == P && GLoad::R == R, "register input must match the stage");
        float2 v[P][R];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int j = t + p * T;
            if (t < T && (S % T == 0 || j < S)) {
                // Element j + r S (0 <= j < S) of the padded area: one base and constant offsets
                // lds_pad(r S) when 16 divides S or S divides 16 (else the index per element).
                const float2* sj = s + (S % 16 == 0 ? lds_pad(j) : j);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if constexpr (REG_IN)
                        v[p][r] = gload.v[p][r];
                    else if constexpr (FIRST)
                        v[p][r] = gload(j + r * S);
                    else if constexpr (S % 16 == 0 || 16 % S == 0)
                        v[p][r] = sj[lds_pad(r * S)];
                    else
                        v[p][r] = s[lds_pad(j + r * S)];
                }
            }
        }
        __syncthreads();  // every read of this stage precedes any write into `s`
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int j = t + p * T;
            if (t < T && (S % T == 0 || j < S)) {
                const int k = j % L;
                if constexpr (L > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r)
                        v[p][r] = cmul(v[p][r], twiddle<DIR>(tw, r * k * (N / (L * R))));
                }
                DFT<R, DIR>::run(v[p]);
                const int base = (j - k) * R + k;
                // Outputs base + r L, where (j - k) R is a multiple of L R and k < L: with 16 | L the
                // padded index is lds_pad(base) + lds_pad(r L); with L | 16 and 16 | L R it is
                // lds_pad((j - k) R) + k + lds_pad(r L); with L R | 16 the block lies in one 16-slot
                // segment: lds_pad((j - k) R) + k + r L.
                constexpr int WR = L % 16 == 0 ? 1 : (16 % L == 0 && (L * R) % 16 == 0) ? 2 : 16 % (L * R) == 0 ? 3 : 0;
                float2* sb = s + (WR == 1 ? lds_pad(base) : lds_pad((j - k) * R) + k);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if constexpr (LAST)
                        gstore(base + r * L, v[p][r]);
                    else if constexpr (WR == 1 || WR == 2)
                        sb[lds_pad(r * L)] = v[p][r];
                    else if constexpr (WR == 3)
                        sb[r * L] = v[p][r];
                    else
                        s[lds_pad(base + r * L)] = v[p][r];
                }
            }
        }
        if constexpr (!LAST) {
            __syncthreads();
            Stockham<N, T, DIR, L * R, Rest...>::run(t, s, tw, gload, gstore);
        }
    }
};

// Workgroup index remapped so that consecutive logical workgroups run on one XCD. The dispatcher
// hands workgroup i to XCD i mod 8; with per = n / 8, logical = (i mod 8) * per + i / 8 for the
// first 8 * per workgroups (a bijection onto [0, 8 * per)), the tail keeps its index. Tiles whose
// rows share 128-byte lines then meet in the same L2 instead of fetching the line once per XCD.
__device__ __forceinline__ int xcd_block(int i, int n)
{
    const int per = n >> 3;
    if (i >= per * 8) return i;
    return (i & 7) * per + (i >> 3);
}

// Radix plans per transform length. T = threads cooperating on one transform.
template <int N>
struct Plan;
template <>
struct Plan<32> {
    static constexpr int T = 4;
    static constexpr int R0 = 8;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<32, T, DIR, 1, 8, 4>::run(t, s, tw, gl, gs);
    }
};
template <>
struct Plan<64> {
    static constexpr int T = 8;
    static constexpr int R0 = 8;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<64, T, DIR, 1, 8, 8>::run(t, s, tw, gl, gs);
    }
};
template <>
struct Plan<128> {
    static constexpr int T = 16;
    static constexpr int R0 = 8;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<128, T, DIR, 1, 8, 16>::run(t, s, tw, gl, gs);
    }
};
template <>
struct Plan<192> {
    static constexpr int T = 16;
    static constexpr int R0 = 12;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<192, T, DIR, 1, 12, 16>::run(t, s, tw, gl, gs);
    }
};
template <>
struct Plan<256> {
    static constexpr int T = 16;
    static constexpr int R0 = 16;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<256, T, DIR, 1, 16, 16>::run(t, s, tw, gl, gs);
    }
};
template <>
struct Plan<384> {
    static constexpr int T = 16;
    static constexpr int R0 = 3;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<384, T, DIR, 1, 3, 8, 16>::run(t, s, tw, gl, gs);
    }
};
template <>
struct Plan<512> {  // three radix-8 stages keep all 64 threads busy (16 x 32 leaves half idle in its second stage)
    static constexpr int T = 64;
    static constexpr int R0 = 8;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<512, T, DIR, 1, 8, 8, 8>::run(t, s, tw, gl, gs);
    }
};
template <>
struct Plan<1024> {
    static constexpr int T = 32;
    static constexpr int R0 = 32;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<1024, T, DIR, 1, 32, 32>::run(t, s, tw, gl, gs);
    }
};

// Alternative radix plan per length with the same thread count, a tuning choice for the volume y pass
// (the autotuner keeps it only where it measures faster): 32 = 4 x 8, 64 = 4 x 4 x 4, 128 = 16 x 8,
// 192 = 16 x 12, 256 = 4 x 4 x 16.
template <int N>
struct PlanAlt;
template <>
struct PlanAlt<32> {
    static constexpr int T = Plan<32>::T;
    static constexpr int R0 = 4;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<32, T, DIR, 1, 4, 8>::run(t, s, tw, gl, gs);
    }
};
template <>
struct PlanAlt<64> {
    static constexpr int T = Plan<64>::T;
    static constexpr int R0 = 4;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<64, T, DIR, 1, 4, 4, 4>::run(t, s, tw, gl, gs);
    }
};
template <>
struct PlanAlt<128> {
    static constexpr int T = Plan<128>::T;
    static constexpr int R0 = 16;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<128, T, DIR, 1, 16, 8>::run(t, s, tw, gl, gs);
    }
};
template <>
struct PlanAlt<192> {
    static constexpr int T = Plan<192>::T;
    static constexpr int R0 = 16;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<192, T, DIR, 1, 16, 12>::run(t, s, tw, gl, gs);
    }
};
template <>
struct PlanAlt<256> {
    static constexpr int T = Plan<256>::T;
    static constexpr int R0 = 4;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<256, T, DIR, 1, 4, 4, 16>::run(t, s, tw, gl, gs);
    }
};
template <>
struct PlanAlt<512> {
    static constexpr int T = Plan<512>::T;
    static constexpr int R0 = 8;  // first radix
    template <int DIR, class TW, class GL, class GS>
    __device__ __forceinline__ static void run(int t, float2* s, const TW& tw, GL& gl, GS& gs)
    {
        Stockham<512, T, DIR, 1, 8, 8, 8>::run(t, s, tw, gl, gs);
    }
};

}  // namespace ctdm
