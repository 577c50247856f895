// Goldstein branch-cut unwrapping for gfx950, bit-exact with the host specification
// (reference cudaGoldsteinUnwrap2D, Goldstein.cu:5-50; residues :52-103, cuts :105-248, flood fill
// :263-641). Integer work throughout: residues int8, cuts uint8, 2*pi counts int32.
//
//   1. k_gs_residues  one thread per pixel: the reference's wrapped differences with its float /
//                     double promotions, summed in float, thresholded at +-6.
//   2. branch cuts, deterministic raster order (the reference runs its search for all residues
//      concurrently, which is racy; the raster order makes it reproducible). Three phases:
//      k_gs_rank + k_gs_cand list every residue's first hits on the initial charges in parallel (one
//      thread per residue), k_gs_commit (one wavefront per map, presence at 1 bit per pixel, in LDS when
//      it fits, else in global memory) walks the residues in order and takes the first listed hit still
//      present, and k_gs_draw draws the chosen cuts in parallel.
//   3. flood fill: level-synchronous breadth-first search over non-cut pixels from a fixed start; each
//                     new pixel takes k from its first neighbour of the previous level in the order
//                     right, left, below, above. Maps whose cut bits fit in LDS (up to 1.3 M pixels):
//                     k_gs_bfs_lev (one workgroup per map, one barrier per level) stores each pixel's
//                     level, k_gs_par_lev takes the parents from the levels, and k_gs_ptr_init /
//                     k_gs_ptr_jump (whole device) sum the steps along the parent chains by pointer
//                     jumping. Larger maps: k_gs_bfs, with k along the levels.
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

// fmodf(a, b) for b > 0, then shifted into [0, b). For |a| < 2b, fmodf is a itself or a - sign(a) b,
// and |a| - b is exact there (b <= |a| <= 2b); other a (wrapped phases far outside [-pi, pi]) take fmodf.
__device__ __forceinline__ float gs_mod(float a, float b)
{
    const float aa = fabsf(a);
    float r = aa < b ? a : (aa < __fmul_rn(2.f, b) ? copysignf(__fsub_rn(aa, b), a) : fmodf(a, b));
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

// Per-pixel kernels run on a grid of (column tiles, rows, maps): column, row and map come from the
// grid, with no division.
__device__ __forceinline__ void pixel_of_grid(int W, int& x, int& y, long long& map)
{
    x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    y = (int)blockIdx.y;
    map = blockIdx.z;
}

// ---------------------------------------------------------------- 1. residues
__global__ void k_gs_residues(const float* __restrict__ phi, int batch, int H, int W, int8_t* __restrict__ res)
{
    const long long n = (long long)H * W;
    int i, j;
    long long b;
    pixel_of_grid(W, i, j, b);
    if (i >= W) return;
    const int p = i + W * j;
    const long long g = b * n + p;
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
// Word w of the bit array of a byte map's nonzero entries (bit b: byte 32 w + b; bytes past n read as 0).
// Two 16-byte loads where the word's bytes sit on a 16-byte boundary, else byte loads at clamped
// addresses, all issued before use.
__device__ __forceinline__ unsigned nonzero_bits(const uint8_t* __restrict__ m, long long n, int w)
{
    const long long p0 = (long long)w * 32;
    unsigned v = 0;
    if (p0 + 32 <= n && ((reinterpret_cast<uintptr_t>(m + p0) & 15) == 0)) {
        const uint4 a = *reinterpret_cast<const uint4*>(m + p0);
        const uint4 b = *reinterpret_cast<const uint4*>(m + p0 + 16);
        const unsigned q[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int e = 0; e < 4; ++e) v |= (unsigned)(((q[k] >> (8 * e)) & 0xffu) != 0) << (4 * k + e);
    } else {
        uint8_t q[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) q[k] = m[min(p0 + k, n - 1)];
#pragma unroll
        for (int k = 0; k < 32; ++k) v |= (unsigned)(p0 + k < n && q[k] != 0) << k;
    }
    return v;
}

// Word w of the bit array of a residue map's negative charges (bit b: pixel 32 w + b; pixels past n: 0).
__device__ __forceinline__ unsigned negative_bits(const int8_t* __restrict__ m, long long n, int w)
{
    const long long p0 = (long long)w * 32;
    int8_t q[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) q[k] = m[min(p0 + k, n - 1)];
    unsigned v = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) v |= (unsigned)(p0 + k < n && q[k] < 0) << k;
    return v;
}

// Charges still present: 1 bit per pixel, 32 pixels per word. A charge keeps its sign until it is
// removed, so the sign is read from the residue map. The words live in LDS when a map's words fit
// (maps up to 1.3 M pixels in 160 KiB), else in global memory; either way one wavefront reads and
// clears them (in global memory through atomics at device scope, which bypass the vector L1).
template <bool LDSW>
struct Presence {
    unsigned* w;
    __device__ __forceinline__ unsigned get(int i) const
    {
        if constexpr (LDSW)
            return w[i];
        else
            return __hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ bool at(int pix) const { return (get(pix >> 5) >> (pix & 31)) & 1u; }
    __device__ __forceinline__ void clear(int pix) const
    {
        const unsigned m = ~(1u << (pix & 31));
        if constexpr (LDSW)
            atomicAnd(w + (pix >> 5), m);
        else
            __hip_atomic_fetch_and(w + (pix >> 5), m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// ---------------------------------------------------------------- 2b. branch cuts in three phases
// The same result as the raster-order search, with most of the work in parallel. Charges only ever
// disappear, and a border point is always a hit, so the first hit of a residue's scan at the time it is
// processed is the first point of its scan that was a hit initially and is still present. Hence:
//   1. every residue (in parallel) lists the first GS_K hits of its scan on the initial charges, stopping
//      at a border point (always valid) or when the box radius runs out;
//   2. one wavefront per map walks the residues in raster order and takes the first listed hit that is
//      still present; only when all GS_K listed charges are gone does it scan again (a full rescan from
//      radius 1 finds the same point as continuing the scan: the earlier points are all non-hits now);
//   3. the chosen cuts are drawn in parallel (cut pixels are only ever set to 1).
constexpr int GS_K = 4;
constexpr int GS_META_SELF = 8;        // residue on the border: balanced by a one-pixel cut
constexpr int GS_META_TERM = 16;       // the list ends with a border point
constexpr int GS_META_EXHAUSTED = 32;  // the scan reached the largest box without GS_K hits or a border
constexpr int GS_META_NEG = 64;        // the residue's charge is -1

// Residues per block of 1024 pixels.
__global__ __launch_bounds__(1024) void k_gs_count(const int8_t* __restrict__ res, long long n, int nblk,
                                                   int* __restrict__ blkcnt)
{
    __shared__ int wc[16];
    const long long map = blockIdx.x / nblk, blk = blockIdx.x % nblk;
    const long long p = blk * 1024 + threadIdx.x;
    const bool q = p < n && res[map * n + p] != 0;
    const unsigned long long m = __ballot(q);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < 16; ++w) t += wc[w];
        blkcnt[blockIdx.x] = t;
    }
}

// Exclusive scan of the block counts of one map (one thread per map; a few hundred blocks).
__global__ void k_gs_scan(int* __restrict__ blkcnt, int nblk, int batch, int* __restrict__ nres)
{
    const int map = blockIdx.x * blockDim.x + threadIdx.x;
    if (map >= batch) return;
    int acc = 0;
    for (int b = 0; b < nblk; ++b) {
        const int v = blkcnt[(long long)map * nblk + b];
        blkcnt[(long long)map * nblk + b] = acc;
        acc += v;
    }
    nres[map] = acc;
}

// Phase 1a: one thread per pixel; each residue writes its pixel at its raster rank (rec_pm.x).
__global__ __launch_bounds__(1024) void k_gs_rank(const int8_t* __restrict__ res, int H, int W, int nblk,
                                                  const int* __restrict__ blkoff, int2* __restrict__ rec_pm)
{
    __shared__ int wc[16];
    const long long n = (long long)H * W;
    const long long map = blockIdx.x / nblk, blk = blockIdx.x % nblk;
    const long long p = blk * 1024 + threadIdx.x;
    const int q = p < n ? res[map * n + p] : 0;
    const unsigned long long m = __ballot(q != 0);
    const int lane = threadIdx.x & 63;
    if (lane == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (q == 0) return;
    int rank = blkoff[blockIdx.x] + __popcll(m & ((1ull << lane) - 1ull));
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) rank += wc[w];
    rec_pm[map * n + rank] = make_int2((int)p, 0);
}

// Phase 1b: one thread per residue (raster rank), so that the scans fill whole wavefronts; the record
// gets the residue's list. A listed hit is encoded as kind << 30 | target pixel (kind 1: column border,
// cut along the row; 2: row border, cut along the column; 3: opposite charge), kind 0 with the
// residue's own pixel for a border residue.
__global__ __launch_bounds__(1024) void k_gs_cand(const int8_t* __restrict__ res, int H, int W, int nblk,
                                                  const int* __restrict__ nres, int2* __restrict__ rec_pm,
                                                  int4* __restrict__ rec_c)
{
    const long long n = (long long)H * W;
    const long long map = blockIdx.x / nblk, rank = (blockIdx.x % nblk) * 1024 + threadIdx.x;
    if (rank >= nres[map]) return;
    const long long o = map * n + rank;
    const int8_t* r = res + map * n;
    const long long p = rec_pm[o].x;
    const int q = r[p];
    const int i = (int)(p % W), j = (int)(p / W);
    int cand[GS_K] = {0, 0, 0, 0};
    int cnt = 0, meta = 0;
    if (i <= 0 || i >= W - 1 || j <= 0 || j >= H - 1) {
        meta = GS_META_SELF;  // its one listed hit is itself (kind 0), which no commit removes
        cand[0] = (int)p;
        cnt = 1;
    } else {
        const int R = min(901, W / 2);
        bool stop = false;
        int rr = 1;
        for (; rr < R && !stop; ++rr) {
            const int m1 = max(i - rr, 0), m2 = min(i + rr, W - 1);
            const int n1 = max(j - rr, 0), n2 = min(j + rr, H - 1);
            for (int mm = m1; mm <= m2 && !stop; ++mm) {
                const int ndel = (mm == m1 || mm == m2) ? 1 : (n2 - n1);
                for (int nn = n1; nn <= n2 && !stop; nn += ndel) {
                    unsigned c = 0;
                    bool hit = true;
                    if (mm <= 0 || mm >= W - 1) {
                        c = (1u << 30) | (unsigned)(mm + W * j);
                        meta |= GS_META_TERM;
                    } else if (nn <= 0 || nn >= H - 1) {
                        c = (2u << 30) | (unsigned)(i + W * nn);
                        meta |= GS_META_TERM;
                    } else if (q * r[mm + (long long)W * nn] == -1) {
                        c = (3u << 30) | (unsigned)(mm + W * nn);
                    } else {
                        hit = false;
                    }
                    if (hit) {
                        cand[cnt++] = (int)c;
                        stop = (meta & GS_META_TERM) || cnt == GS_K;
                    }
                }
            }
        }
        if (!stop) meta |= GS_META_EXHAUSTED;
    }
    rec_pm[o] = make_int2((int)p, meta | cnt | (q < 0 ? GS_META_NEG : 0));
    rec_c[o] = make_int4(cand[0], cand[1], cand[2], cand[3]);
}

__device__ __forceinline__ int sel4(const int4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// Phase 2: one wavefront per map, residues in raster order. Residues are taken 64 at a time, one per
// lane. Each lane reads, once per batch, whether its residue and its listed hits are still present;
// those flags are kept as wave-wide lane masks (one for the residues, one per list position), so a
// residue's choice is a few scalar bit tests, and after each commit the masks lose the lanes whose
// residue or listed hit that commit removed (vector compares of the two removed pixels against every
// lane's pixels). Each committing lane keeps its choice in a register and a mask records which lanes
// committed; the pairs are written (in commit order, which is lane order) and their charges cleared
// once at the end of the batch, so the loop itself issues no memory operation. A residue whose listed charges are all gone clears the batch's
// charges so far (clearing twice is harmless) and scans again on the current charges (the signs of
// the charges it meets come from the residue map).
template <bool LDSW>
__global__ __launch_bounds__(64) void k_gs_commit(const int8_t* __restrict__ res, int H, int W,
                                                  const int* __restrict__ nres, const int2* __restrict__ rec_pm,
                                                  const int4* __restrict__ rec_c, unsigned* __restrict__ gwords,
                                                  int sign_lds, int2* __restrict__ pairs, int* __restrict__ npairs)
{
    extern __shared__ unsigned lwords[];
    const int lane = threadIdx.x;
    const long long n = (long long)H * W;
    const int nw = (int)((n + 31) / 32);
    const Presence<LDSW> pr{LDSW ? lwords : gwords + (long long)blockIdx.x * nw};
    const int8_t* r = res + blockIdx.x * n;
    for (int w = lane; w < nw; w += 64) pr.w[w] = nonzero_bits(reinterpret_cast<const uint8_t*>(r), n, w);
    // sign bits (1: charge -1) after the presence bits when both fit: rescans then read no residue map
    const unsigned* sg = lwords + nw;
    if (sign_lds)
        for (int w = lane; w < nw; w += 64) lwords[nw + w] = negative_bits(r, n, w);
    if constexpr (!LDSW) __threadfence();
    __syncthreads();
    const int total = nres[blockIdx.x];
    const int2* pm = rec_pm + blockIdx.x * n;
    const int4* rc = rec_c + blockIdx.x * n;
    int2* out = pairs + blockIdx.x * n;
    const int R = min(901, W / 2);
    int emitted = 0;
    int2 nxt_pm = lane < total ? pm[lane] : make_int2(0, 0);
    int4 nxt_c = lane < total ? rc[lane] : make_int4(0, 0, 0, 0);
    // The first batch's records are waited for here, so that no batch waits for the records of the one
    // after it (loaded at its start) before its own work.
    __builtin_amdgcn_s_waitcnt(0);
    for (int base = 0; base < total; base += 64) {
        const int2 mypm = nxt_pm;
        const int4 myc = nxt_c;
        const int nidx = base + 64 + lane;  // the next batch's records load while this one commits
        nxt_pm = nidx < total ? pm[nidx] : make_int2(0, 0);
        nxt_c = nidx < total ? rc[nidx] : make_int4(0, 0, 0, 0);
        const bool in = lane < min(64, total - base);
        const int cn = mypm.y & 7;
        // cpix[k]: the pixel of the k-th listed hit when it is a charge (a commit can remove it), else -1
        // (presence is read for every lane and position, at pixel 0 where there is nothing to read, so
        // that the five reads issue together)
        int cpix[GS_K];
        bool hv[GS_K];
#pragma unroll
        for (int k = 0; k < GS_K; ++k) {
            const unsigned c = (unsigned)sel4(myc, k);
            cpix[k] = (c >> 30) == 3u ? (int)(c & 0x3fffffffu) : -1;
            const bool pres = pr.at(max(cpix[k], 0));
            hv[k] = in && k < cn && (cpix[k] < 0 || pres);
        }
        const bool rpres = pr.at(mypm.x);
        unsigned long long live = __ballot(in && rpres);
        unsigned long long h0 = __ballot(hv[0]), h1 = __ballot(hv[1]), h2 = __ballot(hv[2]), h3 = __ballot(hv[3]);
        // this batch's commits: the lanes whose residue committed (commits go in lane order) and, in each
        // such lane, its choice
        unsigned long long done = 0;
        int mych = 0;
        while (live) {
            const int l = __ffsll((long long)live) - 1;
            const unsigned long long bit = 1ull << l;
            live &= ~bit;
            const int p = __builtin_amdgcn_readlane(mypm.x, l);
            // the first listed hit still present (a residue on the border lists itself, kind 0); -1 when
            // there is none (no choice encodes to -1: pixels stay below 2^30 - 1). Scalar selects, no branch.
            const int c0 = __builtin_amdgcn_readlane(myc.x, l), c1 = __builtin_amdgcn_readlane(myc.y, l);
            const int c2 = __builtin_amdgcn_readlane(myc.z, l), c3 = __builtin_amdgcn_readlane(myc.w, l);
            // (instruction order, held by scheduling barriers: the five reads first, so that their
            // latencies overlap, then the vector compares that need only p, then the scalar selects,
            // then the compares that need the choice, and the mask updates last)
            __builtin_amdgcn_sched_barrier(0);
            const unsigned long long e0 = __ballot(cpix[0] == p), e1 = __ballot(cpix[1] == p);
            const unsigned long long e2 = __ballot(cpix[2] == p), e3 = __ballot(cpix[3] == p);
            __builtin_amdgcn_sched_barrier(0);
            // The selects as scalar-ALU assembly: s_and_b64 sets SCC to (mask & bit) != 0, which
            // s_cselect_b32 reads directly (compiled code inserts an s_cmp before each select). Scalar
            // registers only, no memory access; SCC is declared clobbered.
            int chosen;
            unsigned long long tsel;
            asm volatile(
                "s_and_b64 %[t], %[b], %[h3]\n\t"
                "s_cselect_b32 %[ch], %[c3], -1\n\t"
                "s_and_b64 %[t], %[b], %[h2]\n\t"
                "s_cselect_b32 %[ch], %[c2], %[ch]\n\t"
                "s_and_b64 %[t], %[b], %[h1]\n\t"
                "s_cselect_b32 %[ch], %[c1], %[ch]\n\t"
                "s_and_b64 %[t], %[b], %[h0]\n\t"
                "s_cselect_b32 %[ch], %[c0], %[ch]"
                : [ch] "=&s"(chosen), [t] "=&s"(tsel)
                : [b] "s"(bit), [h0] "s"(h0), [h1] "s"(h1), [h2] "s"(h2), [h3] "s"(h3), [c0] "s"(c0), [c1] "s"(c1),
                  [c2] "s"(c2), [c3] "s"(c3)
                : "scc");
            bool found = chosen != -1;
            if (__builtin_expect(!found, 0)) {
                const int meta = __builtin_amdgcn_readlane(mypm.y, l);
                if (meta & (GS_META_TERM | GS_META_EXHAUSTED)) continue;  // nothing within reach: unbalanced
                // every listed charge is gone: scan again on the current charges
                if ((done >> lane) & 1ull) {
                    pr.clear(mypm.x);
                    if (((unsigned)mych >> 30) == 3u) pr.clear(mych & 0x3fffffff);
                }
                if constexpr (!LDSW) __builtin_amdgcn_s_waitcnt(0);
                const int q = (meta & GS_META_NEG) ? -1 : 1;
                const int i = p % W, j = p / W;
                // The box rings 1, 2, ... (radius below R) in scan order, 64 positions per round: each
                // lane takes the next position (ring lr, index lq on that ring), several rings per round
                // while they are short; the first hit in scan order is the lowest lane with a hit. Rings
                // that no border clips have 8 r positions, so a first round 4 or more pixels from every
                // border takes rings 1-3 and the first 16 positions of ring 4 with a fixed lane layout.
                const bool inner = R > 4 && i >= 4 && i <= W - 5 && j >= 4 && j <= H - 5;
                int rc = 1, off0 = 0;  // next position to scan: ring rc, index off0
                while (!found && rc < R) {
                    int lr = R, lq = 0;  // lanes past the last ring scan nothing
                    int start = 0, rr = rc, off = off0;
                    if (rc == 1 && inner) {
                        lr = lane < 8 ? 1 : (lane < 24 ? 2 : (lane < 48 ? 3 : 4));
                        lq = lane - (lane < 8 ? 0 : (lane < 24 ? 8 : (lane < 48 ? 24 : 48)));
                        start = 64;
                        rr = 4;
                        off = 16;
                    }
                    while (start < 64 && rr < R) {
                        const int P = 2 * (min(j + rr, H - 1) - max(j - rr, 0) + 1) +
                                      2 * (min(i + rr, W - 1) - max(i - rr, 0) - 1);
                        const int take = min(P - off, 64 - start);
                        if (lane >= start && lane < start + take) {
                            lr = rr;
                            lq = off + lane - start;
                        }
                        start += take;
                        off += take;
                        if (off == P) {
                            ++rr;
                            off = 0;
                        }
                    }
                    int c = 0;
                    bool hit_here = false;
                    if (lr < R) {
                        // ring lr's positions: its left column top to bottom, then the columns between
                        // alternately on its top and bottom rows, then its right column top to bottom
                        const int m1 = max(i - lr, 0), m2 = min(i + lr, W - 1);
                        const int n1 = max(j - lr, 0), n2 = min(j + lr, H - 1);
                        const int nrow = n2 - n1 + 1, nmid = 2 * (m2 - m1 - 1);
                        const int q2 = lq - nrow;
                        const bool left = lq < nrow, mid = !left && q2 < nmid;
                        const int mm = left ? m1 : (mid ? m1 + 1 + (q2 >> 1) : m2);
                        const int nn = left ? n1 + lq : (mid ? ((q2 & 1) ? n2 : n1) : n1 + (q2 - nmid));
                        const int pp = mm + W * nn;
                        const bool opposite =
                            sign_lds ? ((int)((sg[pp >> 5] >> (pp & 31)) & 1u) == (q > 0)) : (r[pp] == -q);
                        const bool present = pr.at(pp) && opposite;
                        if (mm <= 0 || mm >= W - 1) {
                            c = (int)((1u << 30) | (unsigned)(mm + W * j));
                            hit_here = true;
                        } else if (nn <= 0 || nn >= H - 1) {
                            c = (int)((2u << 30) | (unsigned)(i + W * nn));
                            hit_here = true;
                        } else if (present) {
                            c = (int)((3u << 30) | (unsigned)pp);
                            hit_here = true;
                        }
                    }
                    const unsigned long long hit = __ballot(hit_here);
                    if (hit) {
                        chosen = __builtin_amdgcn_readlane(c, __ffsll((long long)hit) - 1);
                        found = true;
                    }
                    rc = rr;
                    off0 = off;
                }
                if (!found) continue;  // nothing within reach: the residue stays unbalanced
            }
            mych = lane == l ? chosen : mych;
            done |= bit;
            // the masks follow the two removed charges (pixels are >= 0; -2 never matches, and
            // cpix[k] = -1 marks a hit that no commit can remove)
            const int pix = ((unsigned)chosen >> 30) == 3u ? (chosen & 0x3fffffff) : -2;
            const unsigned long long el = __ballot(mypm.x == pix);
            const unsigned long long f0 = __ballot(cpix[0] == pix), f1 = __ballot(cpix[1] == pix);
            const unsigned long long f2 = __ballot(cpix[2] == pix), f3 = __ballot(cpix[3] == pix);
            __builtin_amdgcn_sched_barrier(0);
            live &= ~el;
            h0 &= ~(e0 | f0);
            h1 &= ~(e1 | f1);
            h2 &= ~(e2 | f2);
            h3 &= ~(e3 | f3);
        }
        // the batch's pairs, in commit order, and its charges cleared for the later batches
        if ((done >> lane) & 1ull) {
            const int slot = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(done >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((unsigned)done, 0u));
            out[emitted + slot] = make_int2(mypm.x, mych);
            pr.clear(mypm.x);
            if (((unsigned)mych >> 30) == 3u) pr.clear(mych & 0x3fffffff);
        }
        if constexpr (!LDSW) __builtin_amdgcn_s_waitcnt(0);
        emitted += __popcll(done);
    }
    if (lane == 0) npairs[blockIdx.x] = emitted;
}

// Phase 3: one thread per chosen cut; cut pixels are only set, so the order does not matter.
__global__ void k_gs_draw(const int2* __restrict__ pairs, const int* __restrict__ npairs, int H, int W,
                          uint8_t* __restrict__ cuts)
{
    const long long n = (long long)H * W;
    const long long map = blockIdx.y;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs[map]) return;
    const int2 pr = pairs[map * n + k];
    uint8_t* cut = cuts + map * n;
    const int p = pr.x, kind = (int)((unsigned)pr.y >> 30), t = pr.y & 0x3fffffff;
    if (kind == 0) {
        cut[p] = 1;
        return;
    }
    const int r1 = p % W, c1 = p / W, r2 = t % W, c2 = t / W;
    cut[p] = 1;
    cut[t] = 1;
    const float dR = (float)(r2 - r1), dC = (float)(c2 - c1);
    const float rad = __fsqrt_rn(__fadd_rn(__fmul_rn(dR, dR), __fmul_rn(dC, dC)));
    const float ct = __fdiv_rn(dR, rad), st = __fdiv_rn(dC, rad);
    const int nseg = (int)rintf(rad);
    const int rlo = min(r1, r2), rhi = max(r1, r2), clo = min(c1, c2), chi = max(c1, c2);
    for (int ii = 1; ii <= nseg; ++ii) {
        const int rf = r1 + (int)rintf(__fmul_rn((float)ii, ct));
        const int cf = c1 + (int)rintf(__fmul_rn((float)ii, st));
        if (rf >= rlo && rf <= rhi && cf >= clo && cf <= chi) cut[rf + (long long)W * cf] = 1;
    }
}

// ---------------------------------------------------------------- 3. flood fill (BFS levels)
__device__ __forceinline__ int ld_rel(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
// The same at workgroup scope, for data one workgroup owns until its kernel ends (its waves share a
// CU, so the loads may be served by that CU's cache).
__device__ __forceinline__ int ld_wg(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ __forceinline__ void st_wg(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }

// Queue slot for each calling lane: one atomic per wavefront for the lanes that call it together
// (this module is built without the compiler's atomic optimizer, so the aggregation is explicit).
__device__ __forceinline__ int reserve_slot(int* counter)
{
    const unsigned long long m = __ballot(1);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    int base = 0;
    if (rank == 0) base = atomicAdd(counter, __popcll(m));  // the first active lane
    return __builtin_amdgcn_readfirstlane(base) + rank;
}
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
        st_wg(lev + p, -1);
        st_wg(k + p, INT_MIN);
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
            st_wg(lev + s, 0);
            st_wg(k + s, 0);
            st_wg(cur, s);
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
            const int p = ld_wg(cur + e);
            const int x = p % W, y = p / W;
            for (int t = 0; t < 4; ++t) {
                const int nx = x + dx4[t], ny = y + dy4[t];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                const int q = nx + W * ny;
                if (cut[q]) continue;
                int expected = -1;
                if (__hip_atomic_compare_exchange_strong(lev + q, &expected, d + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP)) {
                    const int pos = reserve_slot(&nnext);
                    st_wg(nxt + pos, q);
                }
            }
        }
        __syncthreads();
        const int nn = nnext;
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const int q = ld_wg(nxt + e);
            const int x = q % W, y = q / W;
            for (int t = 0; t < 4; ++t) {
                const int nx = x + dx4[t], ny = y + dy4[t];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                const int pp = nx + W * ny;
                if (ld_wg(lev + pp) == d) {
                    st_wg(k + q, ld_wg(k + pp) - gs_step2pi(__fsub_rn(ph[q], ph[pp])));
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

// Bit p of a bit mask.
__device__ __forceinline__ bool tbit(const unsigned* b, long long p) { return (b[p >> 5] >> (p & 31)) & 1u; }

// ---------------------------------------------------------------- 3b. flood fill in two stages
// For maps whose cut bits fit in LDS (up to 1.3 M pixels in 160 KiB). Stage 1 (k_gs_bfs_lev, one workgroup
// per map, then k_gs_par_lev over every pixel) finds the levels of the search of k_gs_bfs and each reached
// pixel's parent (the first neighbour of the previous level in the order right, left, below, above) as a
// byte. Stage 2 (k_gs_ptr_init, k_gs_ptr_jump; the whole device) gives every reached pixel the pair
// (parent, -step2pi(phi - phi_parent)) and contracts the parent chains by pointer jumping, (ptr, sum) <-
// (ptr[ptr], sum + sum[ptr]), until each pointer is the start pixel, whose pair is (itself, 0). Then sum
// is the pixel's 2 pi count: the same integers k_gs_bfs adds level by level, in another order (integer
// sums do not depend on it). Unreached pixels keep ptr = -1. Parent codes: 0 unreached, 1..4 right, left,
// below, above, 5 the start pixel.
//
// The search keeps one bit per pixel in LDS, set for cut pixels and for visited ones, so that an atomic
// OR claims a pixel (the returned word tells whether it was still free; a level is one pass and one
// barrier); each claimed pixel's level goes to global memory (stored, never read during the search). The
// parents are taken afterwards from the final levels: levels are the shortest path lengths from the start
// pixel, the same whatever order the search claims pixels in.
__global__ __launch_bounds__(1024) void k_gs_bfs_lev(const uint8_t* __restrict__ cuts, int H, int W, int Q,
                                                     int* __restrict__ level, int* __restrict__ qa,
                                                     int* __restrict__ qb, int* __restrict__ maxdepth)
{
    extern __shared__ unsigned sm[];
    __shared__ int cnt[3];  // entries of levels d, d + 1, d + 2 (mod 3)
    const long long n = (long long)H * W;
    const int nb = (int)((n + 31) / 32);
    unsigned* blk = sm;  // cut or visited
    int* lq[2] = {reinterpret_cast<int*>(sm + nb), reinterpret_cast<int*>(sm + nb) + Q};
    const uint8_t* cut = cuts + blockIdx.x * n;
    int* lev = level + blockIdx.x * n;  // all -1 (unreached) on entry
    int* gq[2] = {qa + blockIdx.x * n, qb + blockIdx.x * n};
    auto qget = [&](int which, int i) -> int { return i < Q ? lq[which][i] : ld_wg(gq[which] + i); };
    auto qput = [&](int which, int i, int v) {
        if (i < Q)
            lq[which][i] = v;
        else
            st_wg(gq[which] + i, v);
    };
    for (int w = threadIdx.x; w < nb; w += blockDim.x) blk[w] = nonzero_bits(cut, n, w);
    __syncthreads();
    if (threadIdx.x == 0) {
        int sx = W / 2, sy = min(20, H - 1);
        while (sx >= 0 && sy < H && tbit(blk, sx + (long long)W * sy)) {
            --sx;
            ++sy;
        }
        cnt[0] = 0;
        if (sx >= 0 && sy < H) {
            const int s0 = sx + W * sy;
            blk[s0 >> 5] |= 1u << (s0 & 31);
            lev[s0] = 0;
            qput(0, 0, (int)((unsigned)sx | ((unsigned)sy << 16)));
            cnt[0] = 1;
        }
        cnt[1] = 0;
        cnt[2] = 0;
    }
    __syncthreads();
    const int dx4[4] = {1, -1, 0, 0}, dy4[4] = {0, 0, 1, -1};
    int d = 0;
    for (;; ++d) {  // one barrier per level
        const int cur = d & 1, nxt = cur ^ 1;                           // queues
        const int c0 = d % 3, c1 = (d + 1) % 3, c2 = (d + 2) % 3;       // counters
        const int nc = cnt[c0];
        if (nc == 0) break;
        if (threadIdx.x == 0) cnt[c2] = 0;  // level d + 2 appends there only after this pass's barrier
        for (int e = threadIdx.x; e < 4 * nc; e += blockDim.x) {  // one (pixel, direction) per thread
            const unsigned v = (unsigned)qget(cur, e >> 2);  // x | y << 16
            const int t = e & 3;
            const int nx = (int)(v & 0xffffu) + dx4[t], ny = (int)(v >> 16) + dy4[t];
            if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
            const int q = nx + W * ny;
            const unsigned bit = 1u << (q & 31);
            if (blk[q >> 5] & bit) continue;
            if (atomicOr(blk + (q >> 5), bit) & bit) continue;  // cut, or claimed by another thread
            lev[q] = d + 1;
            qput(nxt, reserve_slot(&cnt[c1]), (int)((unsigned)nx | ((unsigned)ny << 16)));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && d > 1) atomicMax(maxdepth, d - 1);  // longest parent chain of the map
}

// Parent codes from the levels of k_gs_bfs_lev: a reached pixel of level d + 1 takes its first neighbour
// of level d in the order right, left, below, above (cut and unreached pixels have level -1); the start
// pixel (level 0) gets 5, unreached pixels 0.
__global__ void k_gs_par_lev(const int* __restrict__ level, int batch, int H, int W, uint8_t* __restrict__ par)
{
    const long long n = (long long)H * W;
    int x, y;
    long long b;
    pixel_of_grid(W, x, y, b);
    if (x >= W) return;
    const int p = x + W * y;
    const long long g = b * n + p;
    const int* lev = level + b * n;
    const int d = lev[p];
    uint8_t code = 0;
    if (d == 0) {
        code = 5;
    } else if (d > 0) {
        if (x + 1 < W && lev[p + 1] == d - 1)
            code = 1;
        else if (x > 0 && lev[p - 1] == d - 1)
            code = 2;
        else if (y + 1 < H && lev[p + W] == d - 1)
            code = 3;
        else if (y > 0 && lev[p - W] == d - 1)
            code = 4;
    }
    par[g] = code;
}

__global__ void k_gs_fill_i32(int* __restrict__ p, long long n, int v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Pair (ptr, sum) per pixel, map-local ptr; the start pixel points to itself.
__global__ void k_gs_ptr_init(const float* __restrict__ phi, const uint8_t* __restrict__ par, int batch, int H,
                              int W, int2* __restrict__ pr)
{
    const long long n = (long long)H * W;
    int x, y;
    long long b;
    pixel_of_grid(W, x, y, b);
    if (x >= W) return;
    const int p = x + W * y;
    const long long g = b * n + p;
    const int code = par[g];
    int2 v = make_int2(-1, 0);
    if (code == 5) {
        v = make_int2(p, 0);
    } else if (code >= 1 && code <= 4) {
        const int pp = p + (code == 1 ? 1 : code == 2 ? -1 : code == 3 ? W : -W);
        v = make_int2(pp, -gs_step2pi(__fsub_rn(phi[g], phi[b * n + pp])));
    }
    pr[g] = v;
}

// Pointer jumping, in place: each pixel follows its pointer GS_HOPS times, reading the pairs as they are,
// and keeps the last pointer with the summed steps. Pairs are read and written as single 64-bit accesses,
// so every read sees a consistent pair: an update keeps k(p) = sum(p) + k(ptr(p)) whichever version of a
// pair it reads, and a newer version only shortens the chain. A launch multiplies every pointer's reach by
// at least GS_HOPS + 1 (the pairs read within it may be older than the launch's own updates), so launch
// `step` is needed only while (GS_HOPS + 1)^step is below the longest chain (maxdepth, from the search);
// later launches return at once.
constexpr int GS_HOPS = 3;
__global__ __launch_bounds__(256) void k_gs_ptr_jump(int2* __restrict__ pr, long long n, int step,
                                                     const int* __restrict__ maxdepth)
{
    long long reach = 1;
    for (int i = 0; i < step; ++i) reach *= GS_HOPS + 1;
    if (reach >= (long long)*maxdepth) return;
    unsigned long long* a = reinterpret_cast<unsigned long long*>(pr);
    const long long base = (long long)blockIdx.y * n;  // the grid's y is the map
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const long long g = base + i;
        const unsigned long long mine = __hip_atomic_load(a + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int P = (int)(unsigned)mine;
        if (P < 0) continue;
        int sum = (int)(mine >> 32);
#pragma unroll
        for (int h = 0; h < GS_HOPS; ++h) {  // a reached pixel's pointer is a reached pixel (never negative)
            const unsigned long long up = __hip_atomic_load(a + base + P, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            sum += (int)(up >> 32);
            P = (int)(unsigned)up;
        }
        const unsigned long long nv = (unsigned long long)(unsigned)P | ((unsigned long long)(unsigned)sum << 32);
        if (nv != mine) __hip_atomic_store(a + g, nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------- 4. cut fill + outputs
// k of a pixel: from the fill's int32 counts, or (PTR) from the contracted pairs (ptr < 0: unreached).
template <bool PTR>
__global__ void k_gs_finish(const float* __restrict__ phi, const uint8_t* __restrict__ cuts, const int* __restrict__ kk,
                            const int2* __restrict__ pr, int batch, int H, int W, int* __restrict__ kout,
                            float* __restrict__ unwrapped)
{
    const long long n = (long long)H * W;
    int x, y;
    long long b;
    pixel_of_grid(W, x, y, b);
    if (x >= W) return;
    const int p = x + W * y;
    const long long g = b * n + p;
    const uint8_t* cut = cuts + b * n;
    const float* ph = phi + b * n;
    auto kat = [&](int q) -> int {
        if constexpr (PTR) {
            const int2 v = pr[b * n + q];
            return v.x < 0 ? INT_MIN : v.y;
        } else {
            return kk[b * n + q];
        }
    };
    int kv = kat(p);
    if (cut[p] && x > 0 && x < W - 1 && y > 0 && y < H - 1 &&
        !(cut[p + 1] && cut[p - 1] && cut[p + W] && cut[p - W])) {
        const int nb[4] = {p + 1, p - 1, p + W, p - W};
        for (int t = 0; t < 4; ++t) {
            const int q = nb[t];
            const int kq = kat(q);
            if (!cut[q] && kq != INT_MIN) {  // reached by the fill
                kv = kq - gs_step2pi(__fsub_rn(ph[p], ph[q]));
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
    CTDM_REQUIRE(H <= 65535 && batch <= 65535, "goldstein: at most 65535 rows and 65535 maps per call");
    if (batch == 0) return;
    const long long n = (long long)H * W, tot = n * batch;
    int8_t* r = res ? res : (int8_t*)c.buf("gs.res", tot);
    uint8_t* ct = cuts ? cuts : (uint8_t*)c.buf("gs.cut", tot);
    int* qa = (int*)c.buf("gs.qa", sizeof(int) * tot);
    int* qb = (int*)c.buf("gs.qb", sizeof(int) * tot);
    const dim3 pgrid((unsigned)((W + 255) / 256), (unsigned)H, (unsigned)batch);  // per-pixel kernels
    hipLaunchKernelGGL(k_gs_residues, pgrid, dim3(256), 0, c.stream, wrapped, batch, H, W, r);
    const long long nw = (n + 31) / 32, nb = (n + 31) / 32;
    const size_t lds_max = (size_t)(c.max_lds > 0 ? c.max_lds : 64 * 1024);
    // branch cuts in three phases: candidates in parallel, raster-order commit, drawing
    const size_t commit_lds = sizeof(unsigned) * (size_t)nw;
    const int nblk = (int)((n + 1023) / 1024);
    int* blk = (int*)c.buf("gs.blk", sizeof(int) * (size_t)nblk * batch);
    int* nres = (int*)c.buf("gs.nres", sizeof(int) * (size_t)batch);
    int* npairs = (int*)c.buf("gs.npairs", sizeof(int) * (size_t)batch);
    int2* rec_pm = (int2*)c.buf("gs.recpm", sizeof(int2) * (size_t)tot);
    int4* rec_c = (int4*)c.buf("gs.recc", sizeof(int4) * (size_t)tot);
    int2* pairs = (int2*)c.buf("gs.pairs", sizeof(int2) * (size_t)tot);
    hipLaunchKernelGGL(k_gs_count, dim3((unsigned)(nblk * batch)), dim3(1024), 0, c.stream, r, n, nblk, blk);
    hipLaunchKernelGGL(k_gs_scan, dim3((batch + 63) / 64), dim3(64), 0, c.stream, blk, nblk, batch, nres);
    hipLaunchKernelGGL(k_gs_rank, dim3((unsigned)(nblk * batch)), dim3(1024), 0, c.stream, r, H, W, nblk, blk,
                       rec_pm);
    hipLaunchKernelGGL(k_gs_cand, dim3((unsigned)(nblk * batch)), dim3(1024), 0, c.stream, r, H, W, nblk, nres,
                       rec_pm, rec_c);
    if (commit_lds <= lds_max) {  // presence bits in LDS, and the sign bits beside them when they fit
        const int sign_lds = 2 * commit_lds <= lds_max;
        hipLaunchKernelGGL(k_gs_commit<true>, dim3(batch), dim3(64), (sign_lds ? 2 : 1) * commit_lds, c.stream, r,
                           H, W, nres, rec_pm, rec_c, nullptr, sign_lds, pairs, npairs);
    } else {  // a map's presence bits exceed LDS: they stay in global memory
        unsigned* gw = (unsigned*)c.buf("gs.words", sizeof(unsigned) * (size_t)nw * batch);
        hipLaunchKernelGGL(k_gs_commit<false>, dim3(batch), dim3(64), 0, c.stream, r, H, W, nres, rec_pm, rec_c,
                           gw, 0, pairs, npairs);
    }
    hipLaunchKernelGGL(k_gs_fill_u8, dim3(grid1d(tot, 256)), dim3(256), 0, c.stream, ct, tot);
    hipLaunchKernelGGL(k_gs_draw, dim3((unsigned)((n + 255) / 256), (unsigned)batch), dim3(256), 0, c.stream,
                       pairs, npairs, H, W, ct);
    CTDM_CHECK_LAUNCH("k_gs_draw");
    const size_t bits = sizeof(unsigned) * (size_t)nb;  // cut-or-visited bits
    // the queues get what LDS has left (>= 256 each); their entries pack x and y in 16 bits each
    if (bits + 64 + 2 * sizeof(int) * 256 <= lds_max && W < 65536 && H < 65536) {
        // two-stage fill: levels by the search and parent directions from them, 2 pi counts by pointer jumping
        const int Q = (int)std::min<long long>(n, (long long)((lds_max - 64 - bits) / (2 * sizeof(int))));
        uint8_t* par = (uint8_t*)c.buf("gs.par", (size_t)tot);
        int2* pr = (int2*)c.buf("gs.pr", sizeof(int2) * (size_t)tot);
        int* maxd = (int*)c.buf("gs.maxd", sizeof(int));
        int* lev = (int*)c.buf("gs.level", sizeof(int) * tot);
        hipLaunchKernelGGL(k_gs_fill_u8, dim3(1), dim3(256), 0, c.stream, (uint8_t*)maxd, (long long)sizeof(int));
        hipLaunchKernelGGL(k_gs_fill_i32, dim3(grid1d(tot, 256)), dim3(256), 0, c.stream, lev, tot, -1);
        hipLaunchKernelGGL(k_gs_bfs_lev, dim3(batch), dim3(1024), bits + 2 * sizeof(int) * (size_t)Q, c.stream, ct,
                           H, W, Q, lev, qa, qb, maxd);
        CTDM_CHECK_LAUNCH("k_gs_bfs_lev");
        hipLaunchKernelGGL(k_gs_par_lev, pgrid, dim3(256), 0, c.stream, lev, batch, H, W, par);
        hipLaunchKernelGGL(k_gs_ptr_init, pgrid, dim3(256), 0, c.stream, wrapped, par, batch, H, W, pr);
        // enough launches for any chain in a map ((GS_HOPS + 1)^launches >= n); launches past the longest
        // chain return at once
        int launches = 0;
        for (long long reach = 1; reach < n; reach *= GS_HOPS + 1) ++launches;
        // about 8 workgroups per CU in all, split over the maps (the grid's y)
        const long long per_map = std::max<long long>(1, (long long)std::max(c.cus, 1) * 8 / batch);
        const dim3 jg((unsigned)std::min<long long>(grid1d(n, 256), per_map), (unsigned)batch);
        for (int r = 0; r < launches; ++r)
            hipLaunchKernelGGL(k_gs_ptr_jump, jg, dim3(256), 0, c.stream, pr, n, r, (const int*)maxd);
        CTDM_CHECK_LAUNCH("k_gs_ptr_jump");
        hipLaunchKernelGGL(k_gs_finish<true>, pgrid, dim3(256), 0, c.stream, wrapped, ct, nullptr, pr, batch, H, W,
                           k2pi, unwrapped);
    } else {
        int* kk = (int*)c.buf("gs.k", sizeof(int) * tot);
        int* lev = (int*)c.buf("gs.level", sizeof(int) * tot);
        hipLaunchKernelGGL(k_gs_bfs, dim3(batch), dim3(1024), 0, c.stream, wrapped, ct, H, W, lev, kk, qa, qb);
        CTDM_CHECK_LAUNCH("k_gs_bfs");
        hipLaunchKernelGGL(k_gs_finish<false>, pgrid, dim3(256), 0, c.stream, wrapped, ct, kk, nullptr, batch, H, W,
                           k2pi, unwrapped);
    }
    CTDM_CHECK_LAUNCH("k_gs_finish");
}

}  // namespace ctdm
