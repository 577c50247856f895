// Goldstein branch-cut unwrapping for gfx950, bit-exact with the host specification
// (reference cudaGoldsteinUnwrap2D, Goldstein.cu:5-50; residues :52-103, cuts :105-248, flood fill
// :263-641). Integer work throughout: residues int8, cuts uint8, 2*pi counts int32.
//
//   1. k_gs_residues  one thread per pixel: the reference's wrapped differences with its float /
//                     double promotions, summed in float, thresholded at +-6.
//   2. k_gs_cuts      one wavefront per map: residues in raster order (ballot over packed charge
//                     words), each box radius's perimeter evaluated by 64 lanes at once in the
//                     reference's scan order, the first hit (ballot) decides the cut. The reference
//                     runs this search concurrently for all residues (racy); the raster order makes
//                     it deterministic. Charges live in LDS, 2 bits per pixel.
//   3. k_gs_bfs       one workgroup per map: level-synchronous breadth-first flood fill over
//                     non-cut pixels; after a level is discovered each new pixel takes k from its
//                     first neighbour of the previous level in the order right, left, below, above.
//   4. k_gs_finish    one thread per pixel: interior cut pixels next to a reached pixel take k from
//                     it (same order); unwrapped = phi + 2 pi k, or 0 where unreached.
// Float operations that decide integers are written as single rounded operations and this file is
// built with -ffp-contract=off (the _rn intrinsics are plain operations that the compiler would
// otherwise fuse into FMAs), matching the host specification compiled without contraction.
#include <climits>

#include "internal.hpp"

namespace ctdm {

static constexpr float kTWO_PI_F = 6.28318530717958647692f;
static constexpr float kINV_2PI_F = 0.15915494309189533577f;

__device__ __forceinline__ float gs_mod(float a, float b)
{
    float r = fmodf(a, b);
    if (r < 0.f) r = __fadd_rn(r, b);
    return r;
}

__device__ __forceinline__ float gs_wrap_step(float a, float b)
{
    const float d = __fsub_rn(a, b);
    const float x = (float)__dadd_rn((double)d, M_PI);
    const float m = gs_mod(x, (float)(M_PI * 2.0));
    return (float)__dsub_rn((double)m, M_PI);
}

__device__ __forceinline__ int gs_step2pi(float d) { return (int)rintf(__fmul_rn(d, kINV_2PI_F)); }

// ---------------------------------------------------------------- 1. residues
__global__ void k_gs_residues(const float* __restrict__ phi, int batch, int H, int W, int8_t* __restrict__ res)
{
    const long long n = (long long)H * W;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * batch) return;
    const long long b = g / n;
    const int p = (int)(g % n), i = p % W, j = p / W;
    const float* m = phi + b * n;
    int8_t q = 0;
    if (i < W - 1 && j < H - 1) {
        const float a = m[p], bb = m[p + 1], c = m[p + 1 + W], d = m[p + W];
        const float s = __fadd_rn(__fadd_rn(__fadd_rn(gs_wrap_step(a, bb), gs_wrap_step(bb, c)), gs_wrap_step(c, d)),
                                  gs_wrap_step(d, a));
        q = s >= 6.0f ? 1 : (s <= -6.0f ? -1 : 0);
    }
    res[g] = q;
}

// ---------------------------------------------------------------- 2. branch cuts
// Packed charges: 2 bits per pixel, 16 pixels per word; 0 = none, 1 = +1, 3 = -1.
__device__ __forceinline__ int unpack(unsigned w, int p) { return (int)((w >> ((p & 15) * 2)) & 3u); }

template <bool LDSW>
struct Words {
    unsigned* w;
    __device__ unsigned get(int i) const
    {
        if (LDSW) return w[i];
        return __hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ void clear(int p) const  // caller: one lane
    {
        const unsigned m = ~(3u << ((p & 15) * 2));
        if (LDSW)
            w[p >> 4] &= m;
        else
            __hip_atomic_fetch_and(w + (p >> 4), m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ int charge(int p) const
    {
        const int c = unpack(get(p >> 4), p);
        return c == 1 ? 1 : (c == 3 ? -1 : 0);
    }
};

// cuPlaceBranchCutInternal with signed coordinates; lanes split the fill points.
__device__ void gs_place_cut(uint8_t* cut, int W, int r1, int c1, int r2, int c2, int lane)
{
    if (lane == 0) {
        cut[r1 + (long long)W * c1] = 1;
        cut[r2 + (long long)W * c2] = 1;
    }
    const float dR = (float)(r2 - r1), dC = (float)(c2 - c1);
    const float rad = __fsqrt_rn(__fadd_rn(__fmul_rn(dR, dR), __fmul_rn(dC, dC)));
    const float ct = __fdiv_rn(dR, rad), st = __fdiv_rn(dC, rad);
    const int n = (int)rintf(rad);
    const int rlo = min(r1, r2), rhi = max(r1, r2), clo = min(c1, c2), chi = max(c1, c2);
    for (int ii = 1 + lane; ii <= n; ii += 64) {
        const int rf = r1 + (int)rintf(__fmul_rn((float)ii, ct));
        const int cf = c1 + (int)rintf(__fmul_rn((float)ii, st));
        if (rf >= rlo && rf <= rhi && cf >= clo && cf <= chi) cut[rf + (long long)W * cf] = 1;
    }
}

template <bool LDSW>
__global__ __launch_bounds__(64) void k_gs_cuts(const int8_t* __restrict__ res, int H, int W,
                                                unsigned* __restrict__ gwords, uint8_t* __restrict__ cuts)
{
    extern __shared__ unsigned lwords[];
    const int lane = threadIdx.x;
    const long long n = (long long)H * W;
    const int nw = (int)((n + 15) / 16);
    const int8_t* r = res + blockIdx.x * n;
    uint8_t* cut = cuts + blockIdx.x * n;
    Words<LDSW> ch{LDSW ? lwords : gwords + (long long)blockIdx.x * nw};
    for (int w = lane; w < nw; w += 64) {
        unsigned v = 0;
        for (int k = 0; k < 16; ++k) {
            const long long p = (long long)w * 16 + k;
            const int q = p < n ? r[p] : 0;
            v |= (unsigned)(q == 1 ? 1 : (q == -1 ? 3 : 0)) << (2 * k);
        }
        ch.w[w] = v;
    }
    __syncthreads();
    const int R = min(901, W / 2);
    for (int base = 0; base < nw; base += 64) {
        const int myw = base + lane;
        unsigned long long live = __ballot(myw < nw && ch.get(myw) != 0u);
        while (live) {
            const int l = __ffsll((long long)live) - 1;
            live &= live - 1;
            const int wi = base + l;
            for (int bit = 0; bit < 16; ++bit) {
                const int p = wi * 16 + bit;
                const int q = ch.charge(p);  // re-read: earlier residues may have cleared it
                if (q == 0) continue;
                const int i = p % W, j = p / W;
                if (i <= 0 || i >= W - 1 || j <= 0 || j >= H - 1) {
                    if (lane == 0) {
                        cut[p] = 1;
                        ch.clear(p);
                    }
                    __syncthreads();
                    continue;
                }
                bool done = false;
                for (int rr = 1; rr < R && !done; ++rr) {
                    const int m1 = max(i - rr, 0), m2 = min(i + rr, W - 1);
                    const int n1 = max(j - rr, 0), n2 = min(j + rr, H - 1);
                    const int nrow = n2 - n1 + 1, ncol = m2 - m1 + 1;
                    const int P = 2 * nrow + 2 * (ncol - 2);
                    for (int q0 = 0; q0 < P && !done; q0 += 64) {
                        const int qi = q0 + lane;
                        int kind = 0, mm = 0, nn = 0;
                        if (qi < P) {
                            if (qi < nrow) {
                                mm = m1;
                                nn = n1 + qi;
                            } else if (qi < nrow + 2 * (ncol - 2)) {
                                const int q2 = qi - nrow;
                                mm = m1 + 1 + q2 / 2;
                                nn = (q2 & 1) ? n2 : n1;
                            } else {
                                mm = m2;
                                nn = n1 + (qi - nrow - 2 * (ncol - 2));
                            }
                            if (mm <= 0 || mm >= W - 1)
                                kind = 1;
                            else if (nn <= 0 || nn >= H - 1)
                                kind = 2;
                            else if (q * ch.charge(mm + W * nn) == -1)
                                kind = 3;
                        }
                        const unsigned long long hit = __ballot(kind != 0);
                        if (hit) {
                            const int src = __ffsll((long long)hit) - 1;
                            kind = __shfl(kind, src);
                            mm = __shfl(mm, src);
                            nn = __shfl(nn, src);
                            if (kind == 1)
                                gs_place_cut(cut, W, i, j, mm, j, lane);
                            else if (kind == 2)
                                gs_place_cut(cut, W, i, j, i, nn, lane);
                            else
                                gs_place_cut(cut, W, i, j, mm, nn, lane);
                            if (lane == 0) {
                                ch.clear(p);
                                if (kind == 3) ch.clear(mm + W * nn);
                            }
                            done = true;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

// Whole-map variant for maps whose charges (2 bits) and cut mask (1 bit per pixel) fit in LDS: the cut
// mask is built in LDS and written out once at the end, so the residue loop issues no global stores
// and needs no workgroup barrier (one wavefront: its LDS operations complete in issue order, and every
// branch below is uniform across the wave). The winner of a perimeter scan is read with readlane, and
// charges are cleared with LDS atomics whose result nothing waits for.
__device__ __forceinline__ void cut_bit(unsigned* cw, long long p) { atomicOr(cw + (p >> 5), 1u << (p & 31)); }

__device__ void gs_place_cut_lds(unsigned* cw, int W, int r1, int c1, int r2, int c2, int lane)
{
    if (lane == 0) {
        cut_bit(cw, r1 + (long long)W * c1);
        cut_bit(cw, r2 + (long long)W * c2);
    }
    const float dR = (float)(r2 - r1), dC = (float)(c2 - c1);
    const float rad = __fsqrt_rn(__fadd_rn(__fmul_rn(dR, dR), __fmul_rn(dC, dC)));
    const float ct = __fdiv_rn(dR, rad), st = __fdiv_rn(dC, rad);
    const int n = (int)rintf(rad);
    const int rlo = min(r1, r2), rhi = max(r1, r2), clo = min(c1, c2), chi = max(c1, c2);
    for (int ii = 1 + lane; ii <= n; ii += 64) {
        const int rf = r1 + (int)rintf(__fmul_rn((float)ii, ct));
        const int cf = c1 + (int)rintf(__fmul_rn((float)ii, st));
        if (rf >= rlo && rf <= rhi && cf >= clo && cf <= chi) cut_bit(cw, rf + (long long)W * cf);
    }
}

// Perimeter points of box radii 1, 2 and 3 in the search's scan order (first column top to bottom,
// inner columns top then bottom, last column top to bottom), as (dx, dy): lanes 0-7 ring 1, 8-23
// ring 2, 24-47 ring 3. Evaluating the three rings in one ballot keeps the order of the sequential
// search: the lowest lane with a hit is the first hit of the smallest radius.
struct Rings3 {
    int8_t dx[48], dy[48];
    constexpr Rings3() : dx(), dy()
    {
        int k = 0;
        for (int r = 1; r <= 3; ++r)
            for (int x = -r; x <= r; ++x) {
                if (x == -r || x == r) {
                    for (int y = -r; y <= r; ++y) {
                        dx[k] = (int8_t)x;
                        dy[k] = (int8_t)y;
                        ++k;
                    }
                } else {
                    dx[k] = (int8_t)x;
                    dy[k] = (int8_t)-r;
                    ++k;
                    dx[k] = (int8_t)x;
                    dy[k] = (int8_t)r;
                    ++k;
                }
            }
    }
};
__constant__ constexpr Rings3 kRings3{};

__global__ __launch_bounds__(64) void k_gs_cuts_lds(const int8_t* __restrict__ res, int H, int W,
                                                    uint8_t* __restrict__ cuts)
{
    extern __shared__ unsigned sm[];
    const int lane = threadIdx.x;
    const long long n = (long long)H * W;
    const int nw = (int)((n + 15) / 16), nc = (int)((n + 31) / 32);
    unsigned* chw = sm;       // charges, 2 bits per pixel
    unsigned* cw = sm + nw;   // cut mask, 1 bit per pixel
    const int8_t* r = res + blockIdx.x * n;
    uint8_t* cut = cuts + blockIdx.x * n;
    for (int w = lane; w < nw; w += 64) {
        unsigned v = 0;
        for (int k = 0; k < 16; ++k) {
            const long long p = (long long)w * 16 + k;
            const int q = p < n ? r[p] : 0;
            v |= (unsigned)(q == 1 ? 1 : (q == -1 ? 3 : 0)) << (2 * k);
        }
        chw[w] = v;
    }
    for (int w = lane; w < nc; w += 64) cw[w] = 0u;
    __syncthreads();
    const int R = min(901, W / 2);
    for (int base = 0; base < nw; base += 64) {
        const int myw = base + lane;
        unsigned long long live = __ballot(myw < nw && chw[myw] != 0u);
        while (live) {
            const int l = __ffsll((long long)live) - 1;
            live &= live - 1;
            const int wi = base + l;
            // Only this wave changes the charges, and only through the two clears in take(), so the
            // word is read once and kept current in a register.
            // readfirstlane: the value is the same in every lane; saying so keeps the residue loop scalar
            unsigned wcur = (unsigned)__builtin_amdgcn_readfirstlane((int)chw[wi]);
            const int j0 = (wi * 16) / W, i0 = wi * 16 - j0 * W;  // the word's first pixel
            while (wcur != 0u) {
                wcur = (unsigned)__builtin_amdgcn_readfirstlane((int)wcur);
                const int bit = (__ffs((int)wcur) - 1) >> 1;
                const int c2 = (int)((wcur >> (2 * bit)) & 3u);
                wcur &= ~(3u << (2 * bit));
                const int p = wi * 16 + bit;
                const int q = c2 == 1 ? 1 : -1;
                int i = i0 + bit, j = j0;
                while (i >= W) {
                    i -= W;
                    ++j;
                }
                if (i <= 0 || i >= W - 1 || j <= 0 || j >= H - 1) {
                    if (lane == 0) {
                        cut_bit(cw, p);
                        atomicAnd(chw + wi, ~(3u << (2 * bit)));
                    }
                    continue;
                }
                auto take = [&](int kind, int mm, int nn) {  // wave-uniform winner: cut and balance
                    if (kind == 1)
                        gs_place_cut_lds(cw, W, i, j, mm, j, lane);
                    else if (kind == 2)
                        gs_place_cut_lds(cw, W, i, j, i, nn, lane);
                    else
                        gs_place_cut_lds(cw, W, i, j, mm, nn, lane);
                    if (lane == 0) atomicAnd(chw + wi, ~(3u << (2 * bit)));
                    if (kind == 3) {
                        const int pp = mm + W * nn;
                        if (lane == 0) atomicAnd(chw + (pp >> 4), ~(3u << ((pp & 15) * 2)));
                        if ((pp >> 4) == wi) wcur &= ~(3u << ((pp & 15) * 2));
                    }
                };
                // kind of one perimeter point: 1 column border, 2 row border, 3 opposite charge
                auto classify = [&](int mm, int nn) -> int {
                    if (mm <= 0 || mm >= W - 1) return 1;
                    if (nn <= 0 || nn >= H - 1) return 2;
                    const int pp = mm + W * nn;
                    const int cc = unpack(chw[pp >> 4], pp);
                    return q * (cc == 1 ? 1 : (cc == 3 ? -1 : 0)) == -1 ? 3 : 0;
                };
                bool done = false;
                int rstart = 1;
                if (R > 3 && i >= 3 && i <= W - 4 && j >= 3 && j <= H - 4) {  // radii 1-3 unclamped
                    int kind = 0, mm = 0, nn = 0;
                    if (lane < 48) {
                        mm = i + kRings3.dx[lane];
                        nn = j + kRings3.dy[lane];
                        kind = classify(mm, nn);
                    }
                    const unsigned long long hit = __ballot(kind != 0);
                    if (hit) {
                        const int src = __ffsll((long long)hit) - 1;  // uniform: read the winner's lane
                        take(__builtin_amdgcn_readlane(kind, src), __builtin_amdgcn_readlane(mm, src),
                             __builtin_amdgcn_readlane(nn, src));
                        done = true;
                    }
                    rstart = 4;
                }
                for (int rr = rstart; rr < R && !done; ++rr) {
                    const int m1 = max(i - rr, 0), m2 = min(i + rr, W - 1);
                    const int n1 = max(j - rr, 0), n2 = min(j + rr, H - 1);
                    const int nrow = n2 - n1 + 1, ncol = m2 - m1 + 1;
                    const int P = 2 * nrow + 2 * (ncol - 2);
                    for (int q0 = 0; q0 < P && !done; q0 += 64) {
                        const int qi = q0 + lane;
                        int kind = 0, mm = 0, nn = 0;
                        if (qi < P) {
                            if (qi < nrow) {
                                mm = m1;
                                nn = n1 + qi;
                            } else if (qi < nrow + 2 * (ncol - 2)) {
                                const int q2 = qi - nrow;
                                mm = m1 + 1 + q2 / 2;
                                nn = (q2 & 1) ? n2 : n1;
                            } else {
                                mm = m2;
                                nn = n1 + (qi - nrow - 2 * (ncol - 2));
                            }
                            kind = classify(mm, nn);
                        }
                        const unsigned long long hit = __ballot(kind != 0);
                        if (hit) {
                            const int src = __ffsll((long long)hit) - 1;
                            take(__builtin_amdgcn_readlane(kind, src), __builtin_amdgcn_readlane(mm, src),
                                 __builtin_amdgcn_readlane(nn, src));
                            done = true;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (long long p = lane; p < n; p += 64) cut[p] = (uint8_t)((cw[p >> 5] >> (p & 31)) & 1u);
}

// ---------------------------------------------------------------- 3. flood fill (BFS levels)
__device__ __forceinline__ int ld_rel(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st_rel(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__global__ __launch_bounds__(1024) void k_gs_bfs(const float* __restrict__ phi, const uint8_t* __restrict__ cuts,
                                                 int H, int W, int* __restrict__ level, int* __restrict__ kk,
                                                 int* __restrict__ qa, int* __restrict__ qb)
{
    __shared__ int ncur, nnext;
    const long long n = (long long)H * W;
    const float* ph = phi + blockIdx.x * n;
    const uint8_t* cut = cuts + blockIdx.x * n;
    int* lev = level + blockIdx.x * n;
    int* k = kk + blockIdx.x * n;
    int* cur = qa + blockIdx.x * n;
    int* nxt = qb + blockIdx.x * n;
    for (long long p = threadIdx.x; p < n; p += blockDim.x) {
        st_rel(lev + p, -1);
        st_rel(k + p, INT_MIN);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int sx = W / 2, sy = min(20, H - 1);
        while (sx >= 0 && sy < H && cut[sx + (long long)W * sy]) {
            --sx;
            ++sy;
        }
        ncur = 0;
        if (sx >= 0 && sy < H) {
            const int s = sx + W * sy;
            st_rel(lev + s, 0);
            st_rel(k + s, 0);
            st_rel(cur, s);
            ncur = 1;
        }
        nnext = 0;
    }
    __syncthreads();
    const int dx4[4] = {1, -1, 0, 0}, dy4[4] = {0, 0, 1, -1};
    for (int d = 0;; ++d) {
        const int nc = ncur;
        if (nc == 0) break;
        for (int e = threadIdx.x; e < nc; e += blockDim.x) {
            const int p = ld_rel(cur + e);
            const int x = p % W, y = p / W;
            for (int t = 0; t < 4; ++t) {
                const int nx = x + dx4[t], ny = y + dy4[t];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                const int q = nx + W * ny;
                if (cut[q]) continue;
                int expected = -1;
                if (__hip_atomic_compare_exchange_strong(lev + q, &expected, d + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) {
                    const int pos = atomicAdd(&nnext, 1);
                    st_rel(nxt + pos, q);
                }
            }
        }
        __syncthreads();
        const int nn = nnext;
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const int q = ld_rel(nxt + e);
            const int x = q % W, y = q / W;
            for (int t = 0; t < 4; ++t) {
                const int nx = x + dx4[t], ny = y + dy4[t];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                const int pp = nx + W * ny;
                if (ld_rel(lev + pp) == d) {
                    st_rel(k + q, ld_rel(k + pp) - gs_step2pi(__fsub_rn(ph[q], ph[pp])));
                    break;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            ncur = nn;
            nnext = 0;
        }
        int* t = cur;
        cur = nxt;
        nxt = t;
        __syncthreads();
    }
}

// Whole-map flood fill for maps whose bitmasks fit in LDS (cut, visited, and level membership for two
// alternating levels: 4 bits per pixel). Same levels and parents as k_gs_bfs; the frontier queues and
// k stay in global memory. Per level: (A) expand the current frontier, marking new pixels visited and
// members of level d+1; (B) each new pixel takes k from its first neighbour (right, left, below, above)
// of level d, and the level d-1 marks are cleared from the mask that level d+1 now shares.
__device__ __forceinline__ bool tbit(const unsigned* b, long long p) { return (b[p >> 5] >> (p & 31)) & 1u; }

__global__ __launch_bounds__(1024) void k_gs_bfs_lds(const float* __restrict__ phi, const uint8_t* __restrict__ cuts,
                                                     int H, int W, int* __restrict__ kk, int* __restrict__ qa,
                                                     int* __restrict__ qb, int* __restrict__ qc)
{
    extern __shared__ unsigned sm[];
    __shared__ int ncur, nnext;
    const long long n = (long long)H * W;
    const int nb = (int)((n + 31) / 32);
    unsigned* cutb = sm;
    unsigned* vis = sm + nb;
    unsigned* lv[2] = {sm + 2 * nb, sm + 3 * nb};
    const float* ph = phi + blockIdx.x * n;
    const uint8_t* cut = cuts + blockIdx.x * n;
    int* k = kk + blockIdx.x * n;
    int* prev = qc + blockIdx.x * n;
    int* cur = qa + blockIdx.x * n;
    int* nxt = qb + blockIdx.x * n;
    for (int w = threadIdx.x; w < nb; w += blockDim.x) {
        unsigned v = 0;
        for (int b = 0; b < 32; ++b) {
            const long long p = (long long)w * 32 + b;
            if (p < n && cut[p]) v |= 1u << b;
        }
        cutb[w] = v;
        vis[w] = 0u;
        lv[0][w] = 0u;
        lv[1][w] = 0u;
    }
    for (long long p = threadIdx.x; p < n; p += blockDim.x) st_rel(k + p, INT_MIN);
    __syncthreads();
    if (threadIdx.x == 0) {
        int sx = W / 2, sy = min(20, H - 1);
        while (sx >= 0 && sy < H && tbit(cutb, sx + (long long)W * sy)) {
            --sx;
            ++sy;
        }
        ncur = 0;
        if (sx >= 0 && sy < H) {
            const int s0 = sx + W * sy;
            vis[s0 >> 5] |= 1u << (s0 & 31);
            lv[0][s0 >> 5] |= 1u << (s0 & 31);
            st_rel(k + s0, 0);
            st_rel(cur, s0);
            ncur = 1;
        }
        nnext = 0;
    }
    __syncthreads();
    const int dx4[4] = {1, -1, 0, 0}, dy4[4] = {0, 0, 1, -1};
    int nprev = 0;
    for (int d = 0;; ++d) {
        const int nc = ncur;
        if (nc == 0) break;
        unsigned* lnext = lv[(d + 1) & 1];
        const unsigned* lcur = lv[d & 1];
        for (int e = threadIdx.x; e < nc; e += blockDim.x) {  // (A)
            const int p = ld_rel(cur + e);
            const int x = p % W, y = p / W;
            for (int t = 0; t < 4; ++t) {
                const int nx = x + dx4[t], ny = y + dy4[t];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                const int q = nx + W * ny;
                if (tbit(cutb, q)) continue;
                const unsigned m = 1u << (q & 31);
                if (atomicOr(vis + (q >> 5), m) & m) continue;
                atomicOr(lnext + (q >> 5), m);
                st_rel(nxt + atomicAdd(&nnext, 1), q);
            }
        }
        __syncthreads();
        const int nn = nnext;
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {  // (B)
            const int q = ld_rel(nxt + e);
            const int x = q % W, y = q / W;
            for (int t = 0; t < 4; ++t) {
                const int nx = x + dx4[t], ny = y + dy4[t];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                const int pp = nx + W * ny;
                if (tbit(lcur, pp)) {
                    st_rel(k + q, ld_rel(k + pp) - gs_step2pi(__fsub_rn(ph[q], ph[pp])));
                    break;
                }
            }
        }
        for (int e = threadIdx.x; e < nprev; e += blockDim.x) {  // level d-1 leaves the shared mask
            const int q = ld_rel(prev + e);
            atomicAnd(lnext + (q >> 5), ~(1u << (q & 31)));
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            ncur = nn;
            nnext = 0;
        }
        nprev = nc;
        int* t = prev;
        prev = cur;
        cur = nxt;
        nxt = t;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- 4. cut fill + outputs
__global__ void k_gs_finish(const float* __restrict__ phi, const uint8_t* __restrict__ cuts, const int* __restrict__ kk,
                            int batch, int H, int W, int* __restrict__ kout, float* __restrict__ unwrapped)
{
    const long long n = (long long)H * W;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * batch) return;
    const long long b = g / n;
    const int p = (int)(g % n), x = p % W, y = p / W;
    const uint8_t* cut = cuts + b * n;
    const int* k = kk + b * n;
    const float* ph = phi + b * n;
    int kv = k[p];
    if (cut[p] && x > 0 && x < W - 1 && y > 0 && y < H - 1 &&
        !(cut[p + 1] && cut[p - 1] && cut[p + W] && cut[p - W])) {
        const int nb[4] = {p + 1, p - 1, p + W, p - W};
        for (int t = 0; t < 4; ++t) {
            const int q = nb[t];
            if (!cut[q] && k[q] != INT_MIN) {  // reached by the fill
                kv = k[q] - gs_step2pi(__fsub_rn(ph[p], ph[q]));
                break;
            }
        }
    }
    if (kout) kout[g] = kv;
    unwrapped[g] = kv == INT_MIN ? 0.f : __fadd_rn(ph[p], __fmul_rn((float)kv, kTWO_PI_F));
}

__global__ void k_gs_fill_u8(uint8_t* __restrict__ p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void goldstein(Ctx& c, const float* wrapped, int batch, int H, int W, int8_t* res, uint8_t* cuts, int32_t* k2pi,
               float* unwrapped)
{
    CTDM_REQUIRE(wrapped && unwrapped, "goldstein: null pointer");
    CTDM_REQUIRE(batch >= 0 && H >= 3 && W >= 3, "goldstein: maps must be at least 3 x 3");
    CTDM_REQUIRE((long long)H * W < (1LL << 30), "goldstein: map too large");
    if (batch == 0) return;
    const long long n = (long long)H * W, tot = n * batch;
    int8_t* r = res ? res : (int8_t*)c.buf("gs.res", tot);
    uint8_t* ct = cuts ? cuts : (uint8_t*)c.buf("gs.cut", tot);
    int* kk = (int*)c.buf("gs.k", sizeof(int) * tot);
    int* qa = (int*)c.buf("gs.qa", sizeof(int) * tot);
    int* qb = (int*)c.buf("gs.qb", sizeof(int) * tot);
    hipLaunchKernelGGL(k_gs_residues, dim3(grid1d(tot, 256)), dim3(256), 0, c.stream, wrapped, batch, H, W, r);
    const long long nw = (n + 15) / 16, nb = (n + 31) / 32;
    const size_t lds_max = (size_t)(c.max_lds > 0 ? c.max_lds : 64 * 1024);
    const size_t cuts_lds = sizeof(unsigned) * (size_t)(nw + nb);
    if (cuts_lds <= lds_max) {
        hipLaunchKernelGGL(k_gs_cuts_lds, dim3(batch), dim3(64), cuts_lds, c.stream, r, H, W, ct);
    } else {
        hipLaunchKernelGGL(k_gs_fill_u8, dim3(grid1d(tot, 256)), dim3(256), 0, c.stream, ct, tot);
        const size_t lds = sizeof(unsigned) * (size_t)nw;
        if (lds <= lds_max) {
            hipLaunchKernelGGL((k_gs_cuts<true>), dim3(batch), dim3(64), lds, c.stream, r, H, W, nullptr, ct);
        } else {
            unsigned* gw = (unsigned*)c.buf("gs.words", sizeof(unsigned) * (size_t)nw * batch);
            hipLaunchKernelGGL((k_gs_cuts<false>), dim3(batch), dim3(64), 0, c.stream, r, H, W, gw, ct);
        }
    }
    CTDM_CHECK_LAUNCH("k_gs_cuts");
    const size_t bfs_lds = sizeof(unsigned) * (size_t)(4 * nb);
    if (bfs_lds + 64 <= lds_max) {
        int* qc = (int*)c.buf("gs.qc", sizeof(int) * tot);
        hipLaunchKernelGGL(k_gs_bfs_lds, dim3(batch), dim3(1024), bfs_lds, c.stream, wrapped, ct, H, W, kk, qa, qb, qc);
    } else {
        int* lev = (int*)c.buf("gs.level", sizeof(int) * tot);
        hipLaunchKernelGGL(k_gs_bfs, dim3(batch), dim3(1024), 0, c.stream, wrapped, ct, H, W, lev, kk, qa, qb);
    }
    CTDM_CHECK_LAUNCH("k_gs_bfs");
    hipLaunchKernelGGL(k_gs_finish, dim3(grid1d(tot, 256)), dim3(256), 0, c.stream, wrapped, ct, kk, batch, H, W,
                       k2pi, unwrapped);
    CTDM_CHECK_LAUNCH("k_gs_finish");
}

}  // namespace ctdm
