// Off-axis sideband retrieval ("HilbertTransform", Algorithm A) for gfx950.
//
// Reference sequence (HilbertTransform.cu:1546-1585, 3566-3586, 3618-3639; FFT.cu:220-236, 853-868):
//   FFT2(H x W hologram) -> fftshift -> crop M x M around (row, col) keeping a square passband of
//   half-width `radius` -> ifftshift -> iFFT2 -> x 1/M^2, conjugate -> phase = atan2(BG) - atan2(SP),
//   amp = |SP| / |BG|.
// With S the cropped spectrum re-centred on the sideband (p', q' in [-M/2, M/2)), the field is
//   g(x, y) = sum_{|p'|,|q'| <= radius} F((c0 + p') mod W, (r0 + q') mod H) exp(+2 pi i (p' x + q' y) / M)
// with c0 = col - W/2, r0 = row - H/2 and F the unshifted forward DFT. The batched path evaluates
// exactly this with three kernels per batch of hologram pairs:
//   1. k_row_real: real-input row FFT (W/2-point complex FFT + split), keeps only the passband
//      columns kx = c0 + p'                        -> R[img][y][p'],  img < 2*batch (SP then BG)
//   2. k_col_band: H-point column FFT of each kept column, keeps the passband rows, places them at
//      q' mod M and runs the M-point inverse along y -> G[img][y][p']
//   3. k_row_final: M-point inverse along x for SP and BG rows, then the fused phase / log-amplitude
//      (scale 1/M^2 and the conjugation cancel in arg(SP conj BG) and |SP|/|BG|).
// The fftshifts are index arithmetic; nothing is zero-padded.
#include "fft_dev.hpp"
#include "internal.hpp"

#include <climits>
#include <cmath>
#include <type_traits>

namespace ctdm {

// Two consecutive pixels of a hologram row, read as one complex input element of the row transform.
template <class PX>
struct PxPair;
template <>
struct PxPair<float> {
    using type = float2;
};
template <>
struct PxPair<uint8_t> {
    using type = uchar2;
};
template <>
struct PxPair<uint16_t> {
    using type = ushort2;
};

// Calls f with a null pointer of the pixel type of `px` (the kernels are instantiated per type).
template <class F>
static void with_pixel_type(Px px, F&& f)
{
    switch (px) {
    case Px::F32: f((const float*)nullptr); break;
    case Px::U8: f((const uint8_t*)nullptr); break;
    case Px::U16: f((const uint16_t*)nullptr); break;
    default: fail(CTDM_EINVAL, "unknown hologram pixel format");
    }
}

struct Band {
    int lo;   // first kept offset (p' or q')
    int n;    // number of kept offsets
};

static Band passband(int radius, int M)
{
    const int lo = std::max(-M / 2, -radius);
    const int hi = std::min(M / 2 - 1, radius);
    return Band{lo, hi - lo + 1};
}

// Crop and passband as the kernels read them (same arithmetic as passband()). For any row, column and
// radius >= 0 the band has n <= M entries and every index derived from it is taken modulo H or W, so
// a located crop that later fails validation cannot make a kernel access memory out of bounds.
__host__ __device__ inline CropDev make_crop_dev(int row, int col, int radius, int M, int H, int W)
{
    CropDev d;
    d.row = row;
    d.col = col;
    d.radius = radius;
    d.M = M;
    const int r = radius < 0 ? 0 : radius;
    d.lo = -M / 2 > -r ? -M / 2 : -r;
    const int hi = M / 2 - 1 < r ? M / 2 - 1 : r;
    d.n = hi - d.lo + 1;
    d.k_lo = col - W / 2 + d.lo;
    d.r0 = row - H / 2;
    return d;
}

__global__ void k_set_crop(CropDev* __restrict__ slot, int row, int col, int radius, int M, int H, int W)
{
    *slot = make_crop_dev(row, col, radius, M, H, W);
}

// ---------------------------------------------------------------- 1. real row FFT, pruned output
// Each row of W reals is read as W/2 complex z[n] = x[2n] + i x[2n+1]; Z = FFT_{W/2}(z);
// X[k] = (Z[k] + conj Z[-k]) / 2 - i/2 * exp(-2 pi i k / W) * (Z[k] - conj Z[-k]).
// Output: out[(img*H + y)*nout + q] = X[(k_lo + q) mod W], q < nout. Images img < n0 come from
// src0, the others from src1 (sample holograms first, then backgrounds). PX: the pixel type.
template <int W, int G, class PX>
__global__ __launch_bounds__(G* Plan<W / 2>::T) void k_row_real(
    const PX* __restrict__ src0, int n0, const PX* __restrict__ src1, int nimg, int H,
    const float2* __restrict__ tw_half, const float2* __restrict__ tw_full, int k_lo, int nout,
    const CropDev* __restrict__ cd, float2* __restrict__ out)
{
    constexpr int NH = W / 2;
    if (cd) {  // crop located on the device
        k_lo = cd->k_lo;
        nout = cd->n;
    }
    constexpr int T = Plan<NH>::T;
    __shared__ float2 lds[G * LdsSize<NH>::value];
    __shared__ float2 twl[LdsSize<NH>::value];
    // The G rows of a workgroup lie in one image (H is a multiple of G): image and first row are
    // workgroup-uniform, so the division runs once on the scalar unit, and a workgroup past the last
    // image leaves as a whole before any barrier.
    const int row0 = blockIdx.x * G;  // over nimg * H rows
    const int img_u = row0 / H;
    if (img_u >= nimg) return;
    TwStager<NH, G * T> tws;  // table loads first, its LDS copy after the row's loads are issued
    tws.fetch(tw_half);
    const int g = threadIdx.x / T, t = threadIdx.x % T;
    const long long img = img_u;
    const int y = row0 - img_u * H + g;
    const PX* src = (img < n0 ? src0 + img * (long long)H * W : src1 + (img - n0) * (long long)H * W) +
                    (long long)y * W;
    const auto* src2 = reinterpret_cast<const typename PxPair<PX>::type*>(src);
    float2* s = lds + g * LdsSize<NH>::value;
    // this thread's first QP split twiddles exp(-2 pi i k / W), fetched before the transform to hide
    // their latency (a band of up to QP * T outputs: the whole passband of the batched retrieval);
    // outputs beyond that (the search's full-width rows) load theirs in the loop. The index is in
    // [0, W) for every q, so the reads need no guard (entries for q >= nout go unused).
    constexpr int QN = W / T;  // outputs per thread at most (nout <= W)
    constexpr int QP = QN < 4 ? QN : 4;
    float2 wq[QP];
#pragma unroll
    for (int j = 0; j < QP; ++j) {
        const int q = t + j * T;
        int k = (k_lo + q) % W;
        if (k < 0) k += W;
        wq[j] = tw_full[k];
    }
    StageRegs<NH, T, Plan<NH>::R0> x;
    load_first_stage<NH, T, Plan<NH>::R0>(x, src2, 1, t);
    __builtin_amdgcn_sched_barrier(0);
    tws.put(twl);
    const TwPad twp{twl};
    auto gs = [&](int i, float2 v) { s[lds_pad(i)] = v; };
    Plan<NH>::template run<-1>(t, s, twp, x, gs);
    __syncthreads();
    float2* dst = out + ((long long)img * H + y) * nout;
    auto emit = [&](int q, float2 w) {
        int k = (k_lo + q) % W;
        if (k < 0) k += W;
        const int ka = k % NH;
        const int kb = (NH - ka) % NH;
        const float2 za = s[lds_pad(ka)], zb = s[lds_pad(kb)];
        const float2 e = make_float2(0.5f * (za.x + zb.x), 0.5f * (za.y - zb.y));  // (Z[k] + conj Z[-k]) / 2
        const float2 o = make_float2(0.5f * (za.y + zb.y), -0.5f * (za.x - zb.x)); // (Z[k] - conj Z[-k]) / (2i)
        dst[q] = cadd(e, cmul(w, o));
    };
#pragma unroll
    for (int j = 0; j < QP; ++j)
        if (t + j * T < nout) emit(t + j * T, wq[j]);
    for (int q = t + QP * T; q < nout; q += T) {
        int k = (k_lo + q) % W;
        if (k < 0) k += W;
        emit(q, tw_full[k]);
    }
}

// ---------------------------------------------------------------- 2. column FFT over the band
// 1024-point forward transform of one column by the 32 threads t of that column, as 32 x 32:
// x[t + 32 r] (r < 32) loaded by thread t, a register DFT over r, an exchange through LDS, twiddles
// W_1024^(m t), a second register DFT: thread t ends with X[t + 32 k] in v[k]. The exchange moves
// half of the stage-one outputs at a time (threads t < 16, then t >= 16; each half is two whole
// wavefronts), so the column needs 512 float2 of LDS instead of 1024: with eight columns a
// workgroup takes 35 KiB and four fit on a CU, where the one-pass exchange allowed two.
template <class GL>
__device__ __forceinline__ void col_fft1024_half_lds(int t, float2* s, const float2* __restrict__ tw,
                                                     GL& gl, float2 (&v)[32])
{
#pragma unroll
    for (int r = 0; r < 32; ++r) v[r] = gl(r);
    DFT<32, -1>::run(v);  // v[k] = Y[t][k] = sum_r x[t + 32 r] W_32^(r k)
    float2 u[32];         // u[m] = Y[m][t]
    const int h = t >> 4;
    // Slot 32 m + k of the half holds Y[16 h + m][k]; with t < 32, lds_pad(32 m + k) is
    // 34 m + lds_pad(k), so both sides address one base with constant offsets.
    float2* wr = s + 34 * (t & 15);
    const float2* rd = s + lds_pad(t);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        if (h == half) {
#pragma unroll
            for (int k = 0; k < 32; ++k) wr[lds_pad(k)] = v[k];
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 16; ++m) u[16 * half + m] = rd[34 * m];
        __syncthreads();
    }
    // Twiddles W_1024^(m t) (m t < 1024), read here: the opaque index keeps the reads from being
    // hoisted over the exchange, where their registers would add to its peak.
    int tt = t;
    asm volatile("" : "+v"(tt));
#pragma unroll
    for (int m = 1; m < 32; ++m) u[m] = cmul(u[m], tw[m * tt]);
    DFT<32, -1>::run(u);  // u[k] = X[t + 32 k]
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k] = u[k];
}

// in: R[img][y][q] (H rows of nin columns), column q <-> p' = p_lo + q.
// For each column: F(ky) = FFT_H along y; keep rows ky = (r0 + q') mod H for q' in the band,
// place at index q' mod M of an M-vector (zero elsewhere), inverse M-point FFT ->
// out: G[img][y'][q], y' < M.
template <int H, int M, int C>
__global__ __launch_bounds__(C* Plan<H>::T) void k_col_band(
    const float2* __restrict__ in, int nimg, int nin, int tiles, const float2* __restrict__ tw_h,
    const float2* __restrict__ tw_m, int r0, int q_lo, int q_n, const CropDev* __restrict__ cd,
    float2* __restrict__ out)
{
    constexpr int T = Plan<H>::T;
    constexpr bool HALF = (H == 1024);  // two-pass exchange: the column area holds half a column
    constexpr int LS = LdsStride<HALF ? H / 2 : H, C>::value;
    static_assert(LdsSize<M>::value + (HALF ? T : 0) <= LS,
                  "band transform (and the spare slots) must fit the column area");
    __shared__ float2 lds[C * LS];
    if (cd) {  // crop located on the device; the grid covers the widest band (M columns)
        nin = cd->n;
        r0 = cd->r0;
        q_lo = cd->lo;
        q_n = cd->n;
    }
    const int cidx = threadIdx.x % C, t = threadIdx.x / C;  // lanes walk columns (coalesced)
    const int blk = xcd_block(blockIdx.x, gridDim.x);  // neighbouring column tiles share lines
    if ((blk % tiles) * C >= nin) return;  // a whole tile beyond the band (uniform, before any barrier)
    const int img = blk / tiles;
    const int q = (blk % tiles) * C + cidx;
    const bool valid = q < nin && img < nimg;
    float2* s = lds + cidx * LS;
    if constexpr (HALF) {
        // Element offsets within an image fit 32 bits (H * nin < 2^31); the image base is uniform.
        const float2* base = in + (long long)(img < nimg ? img : 0) * H * nin;
        const unsigned lane = (unsigned)t * nin + (valid ? q : 0);
        auto gl = [&](int r) -> float2 {
            return valid ? base[lane + (unsigned)(32 * r * nin)] : make_float2(0.f, 0.f);
        };
        float2 v[32];
        col_fft1024_half_lds(t, s, tw_h, gl, v);
        // The band rows go to their slots of the M-vector; the last exchange read preceded a barrier.
        // Row ky = t + 32 k is in the band when e = (ky - r0 - q_lo) mod H < q_n; its q' is q_lo + e
        // (the band is shorter than H). Rows outside it go to the thread's spare slot past the
        // M-vector, so that every row is one unconditional store. The opaque copy of t keeps the
        // slots from being formed (and held in registers) before the transform.
        int tt = t;
        asm volatile("" : "+v"(tt));
        const int e0 = tt - r0 - q_lo;
        const int spare = LdsSize<M>::value + tt;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const int e = (e0 + 32 * k) & (H - 1);
            const int qq = q_lo + e;
            s[e < q_n ? lds_pad(qq < 0 ? qq + M : qq) : spare] = v[k];
        }
    } else {
        const float2* col = in + (long long)img * H * nin + (valid ? q : 0);
        auto gl = [&](int i) -> float2 { return valid ? col[(long long)i * nin] : make_float2(0.f, 0.f); };
        auto gs = [&](int i, float2 v) { s[lds_pad(i)] = v; };
        Plan<H>::template run<-1>(t, s, tw_h, gl, gs);
    }
    __syncthreads();
    // Inverse M-point transform along y of the band (ifftshift folded into the index): its input i
    // is the band slot (1024 rows) or the spectrum row ky = (r0 + q') mod H (full column in LDS).
    auto gl2 = [&](int i) -> float2 {
        const int qq = i < M / 2 ? i : i - M;  // q'
        if (qq < q_lo || qq >= q_lo + q_n) return make_float2(0.f, 0.f);
        if constexpr (HALF) return s[lds_pad(i)];
        int ky = (r0 + qq) % H;
        if (ky < 0) ky += H;
        return s[lds_pad(ky)];
    };
    float2* dst = out + (long long)img * M * nin + (valid ? q : 0);
    auto gs2 = [&](int i, float2 v) {
        if (valid) dst[(long long)i * nin] = v;
    };
    Plan<M>::template run<+1>(t, s, tw_m, gl2, gs2);
}

// calcWrapPhase's phase (HilbertTransform.cu:3630): atan2(BG) - atan2(SP) of the conjugated fields, i.e.
// arg SP - arg BG of the fields here, each in (-pi, pi], so the difference lies in (-2 pi, 2 pi) and is
// not wrapped again; non-finite -> 0. Only FastUnwrapping's min-max rescaling (PhaseNormalize) sees the
// difference between this value and the wrapped phase.
__device__ __forceinline__ float calc_wrap_phase(float2 sp, float2 bg)
{
    float p = atan2f(sp.y, sp.x) - atan2f(bg.y, bg.x);
    return isfinite(p) ? p : 0.f;
}

// ---------------------------------------------------------------- 3. final rows + phase / log-amp
// in: G[img][y][q] for img < 2*batch (SP items first). Each workgroup does GR rows of one pair:
// inverse M-point along x of SP and BG rows, then
//   z = g_SP * conj(g_BG): wrapped = atan2(Im z, Re z), logamp = (ln|g_SP|^2 - ln|g_BG|^2) / 2,
//   field = g_SP / g_BG (optional).
template <int M, int GR>
__global__ __launch_bounds__(2 * GR * Plan<M>::T) void k_row_final(
    const float2* __restrict__ in, int batch, int nin, int p_lo, const CropDev* __restrict__ cd,
    const float2* __restrict__ tw_m, float* __restrict__ wrapped, float* __restrict__ logamp,
    float2* __restrict__ field, float* __restrict__ refph)
{
    constexpr int T = Plan<M>::T;
    if (cd) {  // crop located on the device
        nin = cd->n;
        p_lo = cd->lo;
    }
    __shared__ float2 lds[2 * GR * LdsSize<M>::value];
    const int tr = threadIdx.x / T, t = threadIdx.x % T;  // tr in [0, 2*GR): SP rows then BG rows
    const int which = tr / GR, r = tr % GR;
    const int tiles = M / GR;
    const int pair = blockIdx.x / tiles;
    const int y = (blockIdx.x % tiles) * GR + r;
    const long long img = (long long)pair + (long long)which * batch;
    const float2* row = in + (img * M + y) * (long long)nin;
    float2* s = lds + tr * LdsSize<M>::value;
    auto gl = [&](int i) -> float2 {
        const int pp = i < M / 2 ? i : i - M;  // p'
        const int q = pp - p_lo;
        return (q >= 0 && q < nin) ? row[q] : make_float2(0.f, 0.f);
    };
    auto gs = [&](int i, float2 v) { s[lds_pad(i)] = v; };
    Plan<M>::template run<+1>(t, s, tw_m, gl, gs);
    __syncthreads();
    // elementwise combine: GR rows x M pixels, 2*GR*T threads
    for (int e = threadIdx.x; e < GR * M; e += 2 * GR * T) {
        const int rr = e / M, x = e % M;
        const float2 sp = lds[rr * LdsSize<M>::value + lds_pad(x)];
        const float2 bgv = lds[(GR + rr) * LdsSize<M>::value + lds_pad(x)];
        const float2 z = cmulc(sp, bgv);
        const float a2 = sp.x * sp.x + sp.y * sp.y;
        const float b2 = bgv.x * bgv.x + bgv.y * bgv.y;
        float ph = atan2f(z.y, z.x);
        float la = 0.5f * (logf(a2) - logf(b2));
        if (!isfinite(ph)) ph = 0.f;
        if (!isfinite(la)) la = 0.f;
        const long long o = ((long long)pair * M + ((blockIdx.x % tiles) * GR + rr)) * M + x;
        wrapped[o] = ph;
        logamp[o] = la;
        if (field) {
            float2 f = make_float2(z.x / b2, z.y / b2);
            if (!isfinite(f.x) || !isfinite(f.y)) f = make_float2(0.f, 0.f);
            field[o] = f;
        }
        if (refph) refph[o] = calc_wrap_phase(sp, bgv);
    }
}

// ---------------------------------------------------------------- sideband search
// Key = (bits of |F|^2) << 32 | (0xffffffff - raster index): the maximum key is the largest
// magnitude, ties resolved to the first pixel in raster order of the shifted spectrum
// (obtainRadius scans rows j then columns i with a strict '>', HilbertTransform.cu:3439-3450).
__device__ __forceinline__ unsigned long long band_key(float2 v, int i_shifted, int j_shifted, int W)
{
    const float m = v.x * v.x + v.y * v.y;
    return ((unsigned long long)__float_as_uint(m) << 32) |
           (0xffffffffull - (unsigned long long)(i_shifted + (long long)W * j_shifted));
}

// Maximum of a block's per-thread keys (block size a multiple of 64, at most 1024 threads).
__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    __shared__ unsigned long long wmax[16];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = wmax[w] > v ? wmax[w] : v;
    return v;  // valid in thread 0
}

// Column pass of the search fused with the band maximum. The row pass has left the full-width row
// spectrum spec[ky][kx] (unshifted). C columns per workgroup (lanes walk columns, so the loads
// coalesce) get their H-point forward transform along y; every output bin whose shifted row is at or
// beyond row_start enters the key maximum straight from the transform's last stage, and each
// workgroup writes its maximum to part[blockIdx.x] (no spectrum is written back).
template <int H, int C>
__global__ __launch_bounds__(C* Plan<H>::T) void k_col_argmax(const float2* __restrict__ spec, int W, int row_start,
                                                              const float2* __restrict__ tw,
                                                              unsigned long long* __restrict__ part)
{
    constexpr int T = Plan<H>::T;
    constexpr int LS = LdsStride<H, C>::value;
    __shared__ float2 lds[C * LS];
    const int cidx = threadIdx.x % C, t = threadIdx.x / C;
    const int kx = blockIdx.x * C + cidx;  // the grid covers W / C workgroups exactly
    const int i_sh = (kx + W / 2) % W;     // shifted column
    const float2* col = spec + kx;
    float2* s = lds + cidx * LS;
    unsigned long long local = 0;
    auto gl = [&](int ky) -> float2 { return col[(long long)ky * W]; };
    auto gs = [&](int ky, float2 v) {
        const int j = (ky + H / 2) % H;  // shifted row
        if (j >= row_start) {
            const unsigned long long key = band_key(v, i_sh, j, W);
            local = key > local ? key : local;
        }
    };
    Plan<H>::template run<-1>(t, s, tw, gl, gs);
    local = block_max_u64(local);
    if (threadIdx.x == 0) part[blockIdx.x] = local;
}

// Final maximum over the workgroup maxima, written to pinned host memory (one block).
// When `slot` is given, the crop the key stands for (decode_sideband) is written there too.
__global__ __launch_bounds__(256) void k_publish_max(const unsigned long long* __restrict__ part, int n,
                                                     unsigned long long* __restrict__ host,
                                                     CropDev* __restrict__ slot, int H, int W, int M)
{
    unsigned long long v = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v = part[i] > v ? part[i] : v;
    v = block_max_u64(v);
    if (threadIdx.x == 0) {
        if (slot) {
            const long long idx = (long long)(0xffffffffull - (v & 0xffffffffull));
            const int row = (int)(idx / W), col = (int)(idx % W);
            const int dr = row - H / 2;
            *slot = make_crop_dev(row, col, (dr < 0 ? -dr : dr) / 3, M, H, W);
        }
        *host = v;
        __threadfence_system();
    }
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

template <int W, class PX>
static void launch_row_real(Ctx& c, const PX* src0, int n0, const PX* src1, int nimg, int H, int k_lo, int nout,
                            const CropDev* cd, float2* out)
{
    constexpr int G = 256 / Plan<W / 2>::T;
    const long long rows = (long long)nimg * H;
    CTDM_REQUIRE(H % G == 0 && rows <= INT_MAX,
                 "row transform: image height must be a multiple of the row group");
    hipLaunchKernelGGL((k_row_real<W, G, PX>), dim3(grid1d(rows, G)), dim3(G * Plan<W / 2>::T), 0,
                       c.stream, src0, n0, src1, nimg, H, c.tw(W / 2), c.tw(W), k_lo, nout, cd, out);
    CTDM_CHECK_LAUNCH("k_row_real");
}

// cd != nullptr: k_lo and nout are read from the device crop instead. src0 and src1 share one pixel type.
static void row_real(Ctx& c, Pixels src0, int n0, Pixels src1, int nimg, int H, int W, int k_lo, int nout,
                     const CropDev* cd, float2* out)
{
    CTDM_REQUIRE(src0.px == src1.px, "row transform: mixed pixel formats");
    // pixels are read in pairs (one complex element each), so images start at a multiple of two pixels
    const size_t pair = 2 * src0.size();
    CTDM_REQUIRE((uintptr_t)src0.p % pair == 0 && (uintptr_t)src1.p % pair == 0,
                 "hologram pointers must be aligned to two pixels");
    with_pixel_type(src0.px, [&](auto tag) {
        using PX = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
        const PX* a = static_cast<const PX*>(src0.p);
        const PX* b = static_cast<const PX*>(src1.p);
        switch (W) {
        case 256: launch_row_real<256>(c, a, n0, b, nimg, H, k_lo, nout, cd, out); break;
        case 512: launch_row_real<512>(c, a, n0, b, nimg, H, k_lo, nout, cd, out); break;
        case 1024: launch_row_real<1024>(c, a, n0, b, nimg, H, k_lo, nout, cd, out); break;
        default: fail(CTDM_EUNSUPPORTED, "hologram width " + std::to_string(W) + " has no kernel");
        }
    });
}

template <int H, int M>
static void launch_col_band(Ctx& c, const float2* in, int nimg, int nin, int r0, const Band& qb,
                            const CropDev* cd, float2* out)
{
    constexpr int C = 256 / Plan<H>::T;
    const int tiles = (nin + C - 1) / C;  // with a device crop, nin is the widest band (M)
    hipLaunchKernelGGL((k_col_band<H, M, C>), dim3(tiles * nimg), dim3(C * Plan<H>::T), 0, c.stream,
                       in, nimg, nin, tiles, c.tw(H), c.tw(M), r0, qb.lo, qb.n, cd, out);
    CTDM_CHECK_LAUNCH("k_col_band");
}

template <int M>
static void launch_row_final(Ctx& c, const float2* in, int batch, int nin, int p_lo, const CropDev* cd,
                             float* wrapped, float* logamp, float2* field, float* refph)
{
    constexpr int GR = 128 / Plan<M>::T;
    hipLaunchKernelGGL((k_row_final<M, GR>), dim3(batch * (M / GR)), dim3(2 * GR * Plan<M>::T), 0,
                       c.stream, in, batch, nin, p_lo, cd, c.tw(M), wrapped, logamp, field, refph);
    CTDM_CHECK_LAUNCH("k_row_final");
}

template <int H>
static void col_band_m(Ctx& c, int M, const float2* in, int nimg, int nin, int r0, const Band& qb,
                       const CropDev* cd, float2* out)
{
    switch (M) {
    case 32: launch_col_band<H, 32>(c, in, nimg, nin, r0, qb, cd, out); break;
    case 64: launch_col_band<H, 64>(c, in, nimg, nin, r0, qb, cd, out); break;
    case 128: launch_col_band<H, 128>(c, in, nimg, nin, r0, qb, cd, out); break;
    case 192: launch_col_band<H, 192>(c, in, nimg, nin, r0, qb, cd, out); break;
    case 256: launch_col_band<H, 256>(c, in, nimg, nin, r0, qb, cd, out); break;
    default: fail(CTDM_EUNSUPPORTED, "crop size " + std::to_string(M) + " has no kernel");
    }
}

void validate_crop(int H, int W, const ctdm_crop& cr)
{
    CTDM_REQUIRE(H == W, "retrieval: square holograms required");
    CTDM_REQUIRE(cr.M > 0 && cr.M % 2 == 0 && cr.M <= H, "retrieval: bad crop size");
    CTDM_REQUIRE(cr.radius >= 0, "retrieval: negative passband radius");
    CTDM_REQUIRE(cr.row - cr.M / 2 >= 0 && cr.row + cr.M / 2 <= H && cr.col - cr.M / 2 >= 0 &&
                     cr.col + cr.M / 2 <= W,
                 "retrieval: crop window leaves the spectrum");
}

// The three kernels for a band of width n (host crop) or of the device crop cd (buffers and grids then
// sized for the widest band, M).
static void retrieve_core(Ctx& c, Pixels holo, Pixels bg, int batch, int H, int W, int M, const Band& pb,
                          int k_lo, int r0, const CropDev* cd, float* wrapped, float* logamp, float2* field,
                          float* refph)
{
    const int nimg = 2 * batch;
    const int nmax = cd ? M : pb.n;
    float2* R = (float2*)c.buf("ret.R", sizeof(float2) * (size_t)nimg * H * nmax);
    float2* G = (float2*)c.buf("ret.G", sizeof(float2) * (size_t)nimg * M * nmax);
    row_real(c, holo, batch, bg, nimg, H, W, k_lo, pb.n, cd, R);
    switch (H) {
    case 256: col_band_m<256>(c, M, R, nimg, nmax, r0, pb, cd, G); break;
    case 512: col_band_m<512>(c, M, R, nimg, nmax, r0, pb, cd, G); break;
    case 1024: col_band_m<1024>(c, M, R, nimg, nmax, r0, pb, cd, G); break;
    default: fail(CTDM_EUNSUPPORTED, "hologram height " + std::to_string(H) + " has no kernel");
    }
    switch (M) {
    case 32: launch_row_final<32>(c, G, batch, pb.n, pb.lo, cd, wrapped, logamp, field, refph); break;
    case 64: launch_row_final<64>(c, G, batch, pb.n, pb.lo, cd, wrapped, logamp, field, refph); break;
    case 128: launch_row_final<128>(c, G, batch, pb.n, pb.lo, cd, wrapped, logamp, field, refph); break;
    case 192: launch_row_final<192>(c, G, batch, pb.n, pb.lo, cd, wrapped, logamp, field, refph); break;
    case 256: launch_row_final<256>(c, G, batch, pb.n, pb.lo, cd, wrapped, logamp, field, refph); break;
    default: fail(CTDM_EUNSUPPORTED, "crop size has no kernel");
    }
}

void retrieve_batched(Ctx& c, Pixels holo, Pixels bg, int batch, int H, int W, const ctdm_crop& crop,
                      float* wrapped, float* logamp, float2* field)
{
    CTDM_REQUIRE(holo && bg && wrapped && logamp, "retrieve: null pointer");
    CTDM_REQUIRE(batch >= 0, "retrieve: negative batch");
    validate_crop(H, W, crop);
    if (batch == 0) return;
    const Band pb = passband(crop.radius, crop.M);
    retrieve_core(c, holo, bg, batch, H, W, crop.M, pb, crop.col - W / 2 + pb.lo, crop.row - H / 2, nullptr,
                  wrapped, logamp, field, nullptr);
}

void retrieve_batched_dev(Ctx& c, Pixels holo, Pixels bg, int batch, int H, int W, int M, const CropDev* cd,
                          float* wrapped, float* logamp, float2* field, float* refph)
{
    CTDM_REQUIRE(holo && bg && wrapped && logamp && cd, "retrieve: null pointer");
    CTDM_REQUIRE(batch >= 0, "retrieve: negative batch");
    CTDM_REQUIRE(H == W && M > 0 && M % 2 == 0 && M <= H, "retrieval: bad sizes");
    if (batch == 0) return;
    retrieve_core(c, holo, bg, batch, H, W, M, Band{0, M}, 0, 0, cd, wrapped, logamp, field, refph);
}

void set_crop(Ctx& c, CropDev* slot, const ctdm_crop& cr, int H, int W)
{
    hipLaunchKernelGGL(k_set_crop, dim3(1), dim3(1), 0, c.stream, slot, cr.row, cr.col, cr.radius, cr.M, H, W);
    CTDM_CHECK_LAUNCH("k_set_crop");
}

template <int H>
static int launch_col_argmax(Ctx& c, const float2* spec, int W, int row_start, unsigned long long* part)
{
    constexpr int C = 256 / Plan<H>::T;
    static_assert(C <= H, "columns per workgroup");
    const int nblk = W / C;
    CTDM_REQUIRE(W % C == 0 && nblk <= 256, "find_sideband: width has no column tiling");
    hipLaunchKernelGGL((k_col_argmax<H, C>), dim3(nblk), dim3(C * Plan<H>::T), 0, c.stream, spec, W, row_start,
                       c.tw(H), part);
    CTDM_CHECK_LAUNCH("k_col_argmax");
    return nblk;
}

void enqueue_sideband(Ctx& c, Pixels bg, int H, int W, int M, int row_start, CropDev* slot,
                      unsigned long long* key_dev)
{
    CTDM_REQUIRE(bg && key_dev, "find_sideband: null pointer");
    CTDM_REQUIRE(H == W && (H == 256 || H == 512 || H == 1024), "find_sideband: square 256, 512 or 1024 holograms");
    if (row_start <= 0) row_start = 600 * H / 1024;
    CTDM_REQUIRE(row_start < H, "find_sideband: row_start beyond the spectrum");
    float2* spec = (float2*)c.buf("sb.spec", sizeof(float2) * (size_t)H * W);
    // Full-width row spectrum of the real image, then the column transforms fused with the band
    // maximum; the final maximum lands in pinned host memory (and the crop in `slot`).
    row_real(c, bg, 1, bg, 1, H, W, 0, W, nullptr, spec);
    int nblk = 0;
    unsigned long long* part = (unsigned long long*)c.buf("sb.part", sizeof(unsigned long long) * 256);
    switch (H) {
    case 256: nblk = launch_col_argmax<256>(c, spec, W, row_start, part); break;
    case 512: nblk = launch_col_argmax<512>(c, spec, W, row_start, part); break;
    default: nblk = launch_col_argmax<1024>(c, spec, W, row_start, part); break;
    }
    hipLaunchKernelGGL(k_publish_max, dim3(1), dim3(256), 0, c.stream, part, nblk, key_dev, slot, H, W, M);
    CTDM_CHECK_LAUNCH("k_publish_max");
}

ctdm_crop decode_sideband(unsigned long long key, int H, int W, int M)
{
    const long long idx = (long long)(0xffffffffull - (key & 0xffffffffull));
    ctdm_crop out{};
    out.row = (int)(idx / W);
    out.col = (int)(idx % W);
    out.radius = std::abs(out.row - H / 2) / 3;
    out.M = M;
    return out;
}

void find_sideband(Ctx& c, Pixels bg, int H, int W, int M, int row_start, ctdm_crop* out)
{
    CTDM_REQUIRE(bg && out, "find_sideband: null pointer");
    unsigned long long* hb = c.host_words(1);
    unsigned long long* hb_dev = nullptr;
    CTDM_HIP(hipHostGetDevicePointer((void**)&hb_dev, hb, 0));
    enqueue_sideband(c, bg, H, W, M, row_start, nullptr, hb_dev);
    c.spin_sync();
    *out = decode_sideband(*hb, H, W, M);
}

// ---------------------------------------------------------------- illumination angle (AngCal_GPU)
// Order-preserving 32-bit image of a float (negative values below positive ones).
__device__ __forceinline__ unsigned long long ordered_bits(float v)
{
    const unsigned u = __float_as_uint(v);
    return (u & 0x80000000u) ? (unsigned long long)(~u) : (unsigned long long)(u | 0x80000000u);
}

// First maximum of log10|F| (non-finite -> 0, cufftComplex2Real HilbertTransform.cu:2984-2996) over the
// flat fftshifted indices [lo, hi), as thrust::max_element scans (AngCal_GPU :2958-2960).
__global__ void k_argmax_log10(const float2* __restrict__ spec, int H, int W, long long lo, long long hi,
                               unsigned long long* __restrict__ best)
{
    unsigned long long local = 0;
    for (long long g = lo + (long long)blockIdx.x * blockDim.x + threadIdx.x; g < hi;
         g += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(g / W), i = (int)(g % W);
        const float2 v = spec[(long long)((j + H / 2) % H) * W + (i + W / 2) % W];
        float d = log10f(sqrtf(v.x * v.x + v.y * v.y));
        if (!isfinite(d)) d = 0.f;
        const unsigned long long key = (ordered_bits(d) << 32) | (0xffffffffull - (unsigned long long)g);
        local = key > local ? key : local;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(local, off);
        local = o > local ? o : local;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(best, local);
}

void angle_from_image(Ctx& c, Pixels img, int batch, int H, int W, double pixel_um, double mag,
                      double wavelength_nm, double n_med, double* angles)
{
    CTDM_REQUIRE(img && angles, "angle_from_image: null pointer");
    CTDM_REQUIRE(H == W && (H == 256 || H == 512 || H == 1024), "angle_from_image: square 256, 512 or 1024 images");
    CTDM_REQUIRE(pixel_um > 0 && mag > 0 && wavelength_nm > 0 && n_med > 0, "angle_from_image: bad optics");
    const long long N = (long long)H * W;
    float2* spec = (float2*)c.buf("ang.spec", sizeof(float2) * (size_t)N);
    unsigned long long* best = (unsigned long long*)c.buf("ang.best", sizeof(unsigned long long) * (size_t)batch);
    const long long lo = (long long)(N * 0.53), hi = (long long)(N * 0.79);
    std::vector<unsigned long long> h(batch);
    for (int b = 0; b < batch; ++b) {
        const Pixels im = img.offset((size_t)b * N);
        row_real(c, im, 1, im, 1, H, W, 0, W, nullptr, spec);
        fft_cols(c, spec, H, W, W, 1, 0, -1, 1.f, false);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, c.stream, best + b);
        hipLaunchKernelGGL(k_argmax_log10, dim3(256), dim3(256), 0, c.stream, spec, H, W, lo, hi, best + b);
        CTDM_CHECK_LAUNCH("k_argmax_log10");
    }
    CTDM_HIP(hipMemcpyAsync(h.data(), best, sizeof(unsigned long long) * batch, hipMemcpyDeviceToHost, c.stream));
    CTDM_HIP(hipStreamSynchronize(c.stream));
    const double fm = n_med / (wavelength_nm * 1e-9);
    for (int b = 0; b < batch; ++b) {
        const long long idx = (long long)(0xffffffffull - (h[b] & 0xffffffffull));
        const int posX = (int)(idx % W);
        const double fx = (double)(W / 2 - 1 - posX) / ((double)W * pixel_um * 1e-6);  // (511 - posX) / (1024 dx)
        const double st = mag * fx / fm;
        angles[b] = st < 0 ? std::fabs(std::asin(st)) : -std::fabs(std::asin(st));
    }
}

// ---------------------------------------------------------------- row-wise retrieval (Algorithm B)
// extractQPI_AlgorithmB (HilbertTransform.cu:1587-1605) + sequence1DFFT (:1608-1665) + calcWrapPhase:
// each hologram is resampled to H/4 rows by bilinear_interpolation_kernel (:3658-3718; x positions
// i (W-1)/W, y positions j (H-1)/(H/4)), every resampled row is transformed along x, the W/4 bins
// around the sideband column are kept, and an inverse W/4-point transform along x gives the field.
// The reference stores the field transposed (copySharedMem :1704-1726), so map index y' + (H/4) n.
//
// Fused here per workgroup: G resampled rows of one pair (SP and BG), resampling in the load, the
// real row transform as a W/2-point complex transform + split, the band cut, the inverse W/4 transform
// and the phase / log-amplitude combination. With M = W/4 and signed offsets d in [-M/2, M/2), the
// inverse transform's input after the reference's crop and half swap is X[(center + W/2 + d) mod W]
// at position d mod M.

// bilinear_interpolation_kernel's value at output (i, row j) of an H x W image, the reference's
// expression in float (integer pixels widened first).
template <class PX>
__device__ __forceinline__ float resample_b(const PX* __restrict__ img, int H, int W, int i, int iy, float b)
{
    const float xr = (float)(W - 1) / (float)W;
    const int ix = (int)(xr * (float)i);
    const float a = xr * (float)i - (float)ix;
    const PX* r0 = img + (long long)iy * W;
    const float d00 = (float)r0[ix];
    const float d10 = ix + 1 < W ? (float)r0[ix + 1] : 0.f;
    const float d01 = iy + 1 < H ? (float)r0[W + ix] : 0.f;
    const float d11 = (ix + 1 < W && iy + 1 < H) ? (float)r0[W + ix + 1] : 0.f;
    const float t1 = a * d10 + (-d00 * a + d00);
    const float t2 = a * d11 + (-d01 * a + d01);
    return b * t2 + (-t1 * b + t1);
}

// Bin k of the real-input W-point DFT from Z = DFT_{W/2}(x[2m] + i x[2m+1]) held in LDS (padded).
template <int W>
__device__ __forceinline__ float2 real_bin(const float2* s, int k, const float2* __restrict__ tw_full)
{
    constexpr int NH = W / 2;
    const int ka = k % NH;
    const int kb = (NH - ka) % NH;
    const float2 za = s[lds_pad(ka)], zb = s[lds_pad(kb)];
    const float2 e = make_float2(0.5f * (za.x + zb.x), 0.5f * (za.y - zb.y));
    const float2 o = make_float2(0.5f * (za.y + zb.y), -0.5f * (za.x - zb.x));
    return cadd(e, cmul(tw_full[k], o));
}

template <int W, int G, class PX>
__global__ __launch_bounds__(2 * G * Plan<W / 2>::T) void k_rows_b(
    const PX* __restrict__ holo, const PX* __restrict__ bg, int H, int center,
    const float2* __restrict__ tw_half, const float2* __restrict__ tw_full, const float2* __restrict__ tw_m,
    float* __restrict__ wrapped, float* __restrict__ logamp, float* __restrict__ refph)
{
    constexpr int NH = W / 2, T = Plan<NH>::T, M = W / 4;
    constexpr int LS = LdsSize<NH>::value;
    __shared__ float2 lds[2 * G * LS];
    const int N2 = H / 4;
    const int tr = threadIdx.x / T, t = threadIdx.x % T;  // transform: SP rows then BG rows
    const int which = tr / G, r = tr % G;
    const int tiles = N2 / G;
    const int pair = blockIdx.x / tiles;
    const int yq = (blockIdx.x % tiles) * G + r;  // resampled row y'
    const PX* img = (which ? bg : holo) + (long long)pair * H * W;
    const float yr = (float)(H - 1) / (float)N2;
    const int iy = (int)(yr * (float)yq);
    const float bfr = yr * (float)yq - (float)iy;
    float2* s = lds + tr * LS;
    auto gl = [&](int m) -> float2 {
        return make_float2(resample_b(img, H, W, 2 * m, iy, bfr), resample_b(img, H, W, 2 * m + 1, iy, bfr));
    };
    auto gs = [&](int k, float2 v) { s[lds_pad(k)] = v; };
    Plan<NH>::template run<-1>(t, s, tw_half, gl, gs);
    __syncthreads();
    // the band: M bins into registers, then back into this transform's LDS area
    constexpr int PB = (M + T - 1) / T;
    float2 band[PB];
#pragma unroll
    for (int p = 0; p < PB; ++p) {
        const int x = t + p * T;
        if (x < M) {
            const int d = x < M / 2 ? x : x - M;
            int k = (center + W / 2 + d) % W;
            if (k < 0) k += W;
            band[p] = real_bin<W>(s, k, tw_full);
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PB; ++p) {
        const int x = t + p * T;
        if (x < M) s[lds_pad(x)] = band[p];
    }
    __syncthreads();
    auto gl2 = [&](int i) -> float2 { return s[lds_pad(i)]; };
    auto gs2 = [&](int n, float2 v) { s[lds_pad(n)] = v; };
    Plan<M>::template run<+1>(t, s, tw_m, gl2, gs2);
    __syncthreads();
    // combine: lanes walk y' fastest (contiguous in the transposed map)
    const int y0 = (blockIdx.x % tiles) * G;
    for (int e = threadIdx.x; e < G * M; e += 2 * G * T) {
        const int rr = e % G, n = e / G;
        const float2 sp = lds[rr * LS + lds_pad(n)];
        const float2 bv = lds[(G + rr) * LS + lds_pad(n)];
        const float2 z = cmulc(sp, bv);
        float ph = atan2f(z.y, z.x);
        float la = 0.5f * (logf(sp.x * sp.x + sp.y * sp.y) - logf(bv.x * bv.x + bv.y * bv.y));
        if (!isfinite(ph)) ph = 0.f;
        if (!isfinite(la)) la = 0.f;
        const long long o = (long long)pair * M * N2 + (long long)(y0 + rr) + (long long)N2 * n;
        wrapped[o] = ph;
        logamp[o] = la;
        if (refph) refph[o] = calc_wrap_phase(sp, bv);
    }
}

// HistogramFT (HilbertTransform.cu:1728-1748) input: |X| of every shifted column of every resampled row
// of one background, mag[y'][s].
template <int W, int G, class PX>
__global__ __launch_bounds__(G * Plan<W / 2>::T) void k_rows_mag(const PX* __restrict__ img, int H,
                                                                 const float2* __restrict__ tw_half,
                                                                 const float2* __restrict__ tw_full,
                                                                 float* __restrict__ mag)
{
    constexpr int NH = W / 2, T = Plan<NH>::T, LS = LdsSize<NH>::value;
    __shared__ float2 lds[G * LS];
    const int N2 = H / 4;
    const int tr = threadIdx.x / T, t = threadIdx.x % T;
    const int yq = blockIdx.x * G + tr;
    const bool valid = yq < N2;
    const float yr = (float)(H - 1) / (float)N2;
    const int iy = valid ? (int)(yr * (float)yq) : 0;
    const float bfr = valid ? yr * (float)yq - (float)iy : 0.f;
    float2* s = lds + tr * LS;
    auto gl = [&](int m) -> float2 {
        return make_float2(resample_b(img, H, W, 2 * m, iy, bfr), resample_b(img, H, W, 2 * m + 1, iy, bfr));
    };
    auto gs = [&](int k, float2 v) { s[lds_pad(k)] = v; };
    Plan<NH>::template run<-1>(t, s, tw_half, gl, gs);
    __syncthreads();
    if (!valid) return;
    for (int sc = t; sc < W; sc += T) {
        const float2 v = real_bin<W>(s, (sc + W / 2) % W, tw_full);
        mag[(long long)yq * W + sc] = sqrtf(v.x * v.x + v.y * v.y);
    }
}

// Column sums over the resampled rows in row order, then the first maximum over [lo, W) as a 64-bit key
// (value bits, inverted column), as thrust::max_element scans.
__global__ void k_rows_hist_argmax(const float* __restrict__ mag, int N2, int W, int lo,
                                   unsigned long long* __restrict__ best)
{
    const int sc = lo + blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long key = 0;
    if (sc < W) {
        float sum = 0.f;
        for (int j = 0; j < N2; ++j) sum += mag[(long long)j * W + sc];
        key = ((unsigned long long)__float_as_uint(sum) << 32) | (0xffffffffull - (unsigned long long)sc);
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(key, off);
        key = o > key ? o : key;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(best, key);
}

static void check_rows_geometry(int H, int W)
{
    CTDM_REQUIRE(H == W, "row-wise retrieval: square holograms required");
    CTDM_REQUIRE(W == 256 || W == 512 || W == 1024, "row-wise retrieval: holograms of 256, 512 or 1024");
}

int find_sideband_rows(Ctx& c, Pixels bg, int H, int W)
{
    CTDM_REQUIRE(bg, "find_sideband_rows: null pointer");
    check_rows_geometry(H, W);
    const int N2 = H / 4;
    float* mag = (float*)c.buf("sbr.mag", sizeof(float) * (size_t)N2 * W);
    unsigned long long* best = (unsigned long long*)c.buf("sbr.best", sizeof(unsigned long long));
    with_pixel_type(bg.px, [&](auto ptag) {
        using PX = std::remove_const_t<std::remove_pointer_t<decltype(ptag)>>;
        const PX* img = static_cast<const PX*>(bg.p);
        auto launch = [&](auto wtag) {
            constexpr int Wc = decltype(wtag)::value;
            constexpr int G = 256 / Plan<Wc / 2>::T;
            hipLaunchKernelGGL((k_rows_mag<Wc, G, PX>), dim3((N2 + G - 1) / G), dim3(G * Plan<Wc / 2>::T), 0,
                               c.stream, img, H, c.tw(Wc / 2), c.tw(Wc), mag);
        };
        switch (W) {
        case 256: launch(std::integral_constant<int, 256>{}); break;
        case 512: launch(std::integral_constant<int, 512>{}); break;
        default: launch(std::integral_constant<int, 1024>{}); break;
        }
    });
    CTDM_CHECK_LAUNCH("k_rows_mag");
    const int lo = (int)(W * 0.6);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, c.stream, best);
    hipLaunchKernelGGL(k_rows_hist_argmax, dim3((W - lo + 255) / 256), dim3(256), 0, c.stream, mag, N2, W, lo, best);
    CTDM_CHECK_LAUNCH("k_rows_hist_argmax");
    unsigned long long h = 0;
    CTDM_HIP(hipMemcpyAsync(&h, best, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    CTDM_HIP(hipStreamSynchronize(c.stream));
    int idx = (int)(0xffffffffull - (h & 0xffffffffull));
    if (W - idx < (W / 4) / 2) idx = W - idx;  // HilbertTransform.cu:1641
    return idx;
}

void retrieve_rows(Ctx& c, Pixels holo, Pixels bg, int batch, int H, int W, int center, float* wrapped,
                   float* logamp, float* refph)
{
    CTDM_REQUIRE(holo && bg && wrapped && logamp, "retrieve_rows: null pointer");
    CTDM_REQUIRE(holo.px == bg.px, "retrieve_rows: mixed pixel formats");
    CTDM_REQUIRE(batch >= 0, "retrieve_rows: negative batch");
    check_rows_geometry(H, W);
    CTDM_REQUIRE(center - W / 8 >= 0 && center + W / 8 <= W,
                 "retrieve_rows: the W/4 band around the sideband column leaves the spectrum");
    if (batch == 0) return;
    with_pixel_type(holo.px, [&](auto ptag) {
        using PX = std::remove_const_t<std::remove_pointer_t<decltype(ptag)>>;
        const PX* sp = static_cast<const PX*>(holo.p);
        const PX* bk = static_cast<const PX*>(bg.p);
        auto launch = [&](auto wtag) {
            constexpr int Wc = decltype(wtag)::value;
            constexpr int G = 128 / Plan<Wc / 2>::T;
            const int N2 = H / 4;
            hipLaunchKernelGGL((k_rows_b<Wc, G, PX>), dim3(batch * (N2 / G)), dim3(2 * G * Plan<Wc / 2>::T), 0,
                               c.stream, sp, bk, H, center, c.tw(Wc / 2), c.tw(Wc), c.tw(Wc / 4), wrapped, logamp,
                               refph);
        };
        switch (W) {
        case 256: launch(std::integral_constant<int, 256>{}); break;
        case 512: launch(std::integral_constant<int, 512>{}); break;
        default: launch(std::integral_constant<int, 1024>{}); break;
        }
    });
    CTDM_CHECK_LAUNCH("k_rows_b");
}

// ---------------------------------------------------------------- plain (reference-structured) path
template <class PX>
__global__ void k_real_to_complex(const PX* __restrict__ in, float2* __restrict__ out, long long n)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) out[g] = make_float2((float)in[g], 0.f);
}

// get1stOrder_new (HilbertTransform.cu:3566-3586), M-generalised.
__global__ void k_get_first_order(float2* __restrict__ out, const float2* __restrict__ in, int radius,
                                  int r, int c, int W, int M)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= M || j >= M) return;
    const int ii = i + (c - M / 2), jj = j + (r - M / 2);
    if (ii >= c - radius && ii <= c + radius && jj >= r - radius && jj <= r + radius)
        out[i + M * j] = in[ii + (long long)W * jj];
    else
        out[i + M * j] = make_float2(0.f, 0.f);
}

// calcWrapPhase (HilbertTransform.cu:3618-3639) on conjugated fields.
__global__ void k_calc_wrap_phase(float* __restrict__ phase, float* __restrict__ amp,
                                  const float2* __restrict__ sp, const float2* __restrict__ bgf, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p = atan2f(bgf[i].y, bgf[i].x) - atan2f(sp[i].y, sp[i].x);
    if (!isfinite(p)) p = 0.f;
    float a = sqrtf(sp[i].y * sp[i].y + sp[i].x * sp[i].x) / sqrtf(bgf[i].y * bgf[i].y + bgf[i].x * bgf[i].x);
    if (!isfinite(a)) a = 0.f;
    phase[i] = p;
    amp[i] = a;
}

void retrieve_plain_one(Ctx& c, Pixels holo, Pixels bg, int H, int W, const ctdm_crop& crop, float* phase,
                        float* amp)
{
    validate_crop(H, W, crop);
    CTDM_REQUIRE(holo.px == bg.px, "retrieval: mixed pixel formats");
    const int M = crop.M;
    const long long n = (long long)H * W;
    float2* SP = (float2*)c.buf("plain.SP", sizeof(float2) * n);
    float2* BG = (float2*)c.buf("plain.BG", sizeof(float2) * n);
    float2* SP2 = (float2*)c.buf("plain.SP2", sizeof(float2) * M * M);
    float2* BG2 = (float2*)c.buf("plain.BG2", sizeof(float2) * M * M);
    with_pixel_type(holo.px, [&](auto ptag) {  // the reference widens its 8-bit frames on the host
        using PX = std::remove_const_t<std::remove_pointer_t<decltype(ptag)>>;
        hipLaunchKernelGGL((k_real_to_complex<PX>), dim3(grid1d(n, 256)), dim3(256), 0, c.stream,
                           static_cast<const PX*>(holo.p), SP, n);
        hipLaunchKernelGGL((k_real_to_complex<PX>), dim3(grid1d(n, 256)), dim3(256), 0, c.stream,
                           static_cast<const PX*>(bg.p), BG, n);
    });
    CTDM_CHECK_LAUNCH("k_real_to_complex");
    for (float2* X : {SP, BG}) {  // cufftExecC2C 2-D forward (HilbertTransform.cu:1564-1565)
        fft_rows(c, X, W, H, -1, 1.f, false);
        fft_cols(c, X, H, W, W, 1, 0, -1, 1.f, false);
    }
    fftshift2d(c, SP, W, H, 1);
    fftshift2d(c, BG, W, H, 1);
    dim3 b2(32, 8), g2((M + 31) / 32, (M + 7) / 8);
    hipLaunchKernelGGL(k_get_first_order, g2, b2, 0, c.stream, SP2, SP, crop.radius, crop.row, crop.col, W, M);
    hipLaunchKernelGGL(k_get_first_order, g2, b2, 0, c.stream, BG2, BG, crop.radius, crop.row, crop.col, W, M);
    CTDM_CHECK_LAUNCH("k_get_first_order");
    fftshift2d(c, SP2, M, M, 1);
    fftshift2d(c, BG2, M, M, 1);
    const float sc = 1.f / ((float)M * (float)M);
    for (float2* X : {SP2, BG2}) {  // cufftExecC2C inverse + scaleFFT2D (scale and conjugate)
        fft_rows(c, X, M, M, +1, 1.f, false);
        fft_cols(c, X, M, M, M, 1, 0, +1, sc, true);
    }
    hipLaunchKernelGGL(k_calc_wrap_phase, dim3(grid1d(M * M, 256)), dim3(256), 0, c.stream, phase, amp, SP2, BG2, M * M);
    CTDM_CHECK_LAUNCH("k_calc_wrap_phase");
}

}  // namespace ctdm
