// Tomographic reconstruction for gfx950: Rytov spectra, Ewald-cap direct mapping into a V^3
// complex64 k-space, 3-D FFTs and the positivity-constrained loop (reference BatchRecon,
// HilbertTransform.cu:1825-2047, and Recon_kernel.cu:144-412).
//
// Geometry (EdwardSphere / fillEdwardSphere, Recon_kernel.cu:171-229): for angle (ang_x, ang_y)
//   f_x0 = f_m0 sin(ang_y), f_y0 = f_m0 sin(ang_x), f_z0 = sqrt(f_m0^2 - f_x0^2 - f_y0^2)
// and spectrum pixel (Nx, Ny) with signed frequencies ii, jj in [-V/2, V/2):
//   f_z = sqrt(f_m0^2 - (ii df + f_x0)^2 - (jj df + f_y0)^2),  Nz = mod(round(( f_z - f_z0) / df), V)
//   F[Nx, Ny, Nz] += 4 pi i f_z spec[Nx, Ny],  hits[...] += 1   (skipped where f_z is not real)
// The voxel index is computed in FP64 with explicitly rounded operations (no contraction), exactly
// as the oracle computes it, so both assign every sample to the same voxel.
//
// Loop (HilbertTransform.cu:1940-1972): n = est_n(f); repeat { f = modify(n); F = FFT3(f);
// F = F_meas on measured voxels; f = iFFT3(F); n = est_n(f) }, then the octant shift (:1974).
// The fused schedule runs each iteration as 4 pencil passes over the volume:
//   X: inverse x-FFT, 1/V^3, P = modify o est_n, forward x-FFT  (one read + one write)
//   Y: forward y-FFT
//   Z: forward z-FFT, data consistency from a compact per-pencil list, inverse z-FFT
//   Y: inverse y-FFT
// The unfused schedule (plain path) launches the reference's kernel sequence: modify, 3 FFT
// passes, check, 3 FFT passes, scale, est_n.
#include "fft_dev.hpp"
#include "internal.hpp"

#include <hip/hip_ext.h>

#include <cmath>
#include <cstring>

namespace ctdm {

// ---------------------------------------------------------------- geometry (host)
EwaldGeom ewald_geom(const ctdm_optics& o, int V)
{
    EwaldGeom g;
    g.fm0 = o.n_med / o.wavelength;
    g.fm02 = g.fm0 * g.fm0;
    g.dx = o.dx;
    g.df = 1.0 / ((double)V * o.dx);
    const double k = 2.0 * M_PI * o.n_med / o.wavelength;
    g.k2 = k * k;
    g.n_med = o.n_med;
    return g;
}

std::vector<AngleRow> angle_rows(const EwaldGeom& g, const ctdm_angle* ang, int A)
{
    CTDM_REQUIRE(ang || A == 0, "angles: null pointer");
    std::vector<AngleRow> r(A);
    for (int a = 0; a < A; ++a) {
        const double fx0 = g.fm0 * std::sin(ang[a].ang_y);  // X/Y swap of Recon_kernel.cu:175-176
        const double fy0 = g.fm0 * std::sin(ang[a].ang_x);
        const double s0 = g.fm02 - fx0 * fx0 - fy0 * fy0;
        r[a] = AngleRow{fx0, fy0, std::sqrt(s0 > 0 ? s0 : 0.0)};
    }
    return r;
}

struct Cap {
    int nz;
    float w;  // 4 pi f_z
    bool ok;
};

// Plain operators under `contract(off)`: the FP64 expressions are rounded exactly as written (the
// host oracle is compiled with -ffp-contract=off), so both pick the same voxel on rounding ties.
__device__ __forceinline__ Cap ewald_cap(int ii, int jj, double fx0, double fy0, double fz0, double fm02,
                                         double df, int V)
{
#pragma clang fp contract(off)
    const double fx = (double)ii * df + fx0;
    const double fy = (double)jj * df + fy0;
    const double s = fm02 - (fx * fx + fy * fy);
    const double s0 = fm02 - fx0 * fx0 - fy0 * fy0;
    Cap c;
    c.ok = !(s0 < 0.0 || s < 0.0);
    const double fz = __builtin_sqrt(s > 0.0 ? s : 0.0);
    const double q = (fz - fz0) / df;
    const double nz = q > 0.0 ? floor(q + 0.5) : ceil(q - 0.5);
    int m = (int)nz % V;
    if (m < 0) m += V;
    c.nz = m;
    c.w = (float)(4.0 * M_PI * fz);
    return c;
}

__device__ __forceinline__ int signed_freq(int n, int V) { return n < V / 2 ? n : n - V; }

// LDS chunk of per-pencil samples staged at a time by the gridding and fit kernels (one bit each in
// a 32-bit owner mask). 32 keeps the gridding pass's LDS at 7 workgroups per CU for 256^3.
constexpr int GRID_CHUNK = 32;
static_assert(GRID_CHUNK <= 32, "owner masks are 32-bit");

// Owner masks of a staged chunk. Thread t of a pencil's T threads owns the z indices z = t + T j;
// interleaved ownership spreads a cap's neighbouring z values over the threads, so the per-thread
// serial accumulation stays short. Whoever stages entry k (z in [0, V)) sets bit k of its owner's
// 32-bit word with an LDS atomic OR, which is order-independent; after the barrier the owner reads its
// word and walks the set bits in increasing k, i.e. in angle (or list) order, so sums are
// deterministic. The owner clears its word before the next chunk is staged.
template <int T>
__device__ __forceinline__ void mark_owner(unsigned* masks, int z, int k)
{
    static_assert((T & (T - 1)) == 0, "owner index needs a power-of-two thread count");
    atomicOr(&masks[z & (T - 1)], 1u << k);
}

void write_rows(Ctx& c, double* dst, const std::vector<AngleRow>& rows, bool cached)
{
    CTDM_REQUIRE(!c.capturing, "angle rows written during graph capture");
    if (rows.empty()) return;
    const size_t n = 3 * rows.size();
    if (cached) {
        std::vector<double> flat(n);
        std::memcpy(flat.data(), rows.data(), sizeof(double) * n);
        auto it = c.table_contents.find(dst);
        if (it != c.table_contents.end() && it->second == flat) return;
        c.table_contents[dst] = std::move(flat);
    }
    CTDM_HIP(hipMemcpyAsync(dst, rows.data(), sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
}

// Device copy of the per-angle rows (fx0, fy0, fz0) in scratch buffer `name` (eager entry points only).
static const double* upload_rows(Ctx& c, const std::string& name, const std::vector<AngleRow>& rows)
{
    double* d = (double*)c.buf(name, sizeof(double) * 3 * std::max<size_t>(rows.size(), 1));
    write_rows(c, d, rows);
    return d;
}

// ---------------------------------------------------------------- Rytov spectrum (batched)
// u(x, y) = logamp + i phi. Reference: u_sp[y' = x][x' = y] (cuCopyArray, transposed), spatial
// fftshift, forward 2-D DFT. With U = DFT2(u): spec[kx' + M ky'] = (-1)^(kx'+ky') U(kx = ky', ky = kx').
// Pass 1 (rows of u): P[a][y][kx] = DFT_x u. Pass 2 (columns kx): DFT_y, store at kx*M + ky.
template <int M, int G>
__global__ __launch_bounds__(G* Plan<M>::T) void k_rytov_rows(const float* __restrict__ phi,
                                                              const float* __restrict__ la, int A,
                                                              const float2* __restrict__ tw,
                                                              float2* __restrict__ P)
{
    constexpr int T = Plan<M>::T;
    __shared__ float2 lds[G * LdsSize<M>::value];
    const int g = threadIdx.x / T, t = threadIdx.x % T;
    const long long row = (long long)blockIdx.x * G + g;
    const bool valid = row < (long long)A * M;
    const long long r = valid ? row : 0;
    float2* s = lds + g * LdsSize<M>::value;
    auto gl = [&](int x) -> float2 { return make_float2(la[r * M + x], phi[r * M + x]); };
    auto gs = [&](int k, float2 v) {
        if (valid) P[r * M + k] = v;
    };
    Plan<M>::template run<-1>(t, s, tw, gl, gs);
}

template <int M, int B>
__global__ __launch_bounds__(B* Plan<M>::T) void k_rytov_cols(const float2* __restrict__ P, int A,
                                                              const float2* __restrict__ tw,
                                                              float2* __restrict__ spec)
{
    constexpr int T = Plan<M>::T;
    constexpr int LS = LdsStride<M, B>::value;
    __shared__ float2 lds[B * LS];
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    const int tiles = M / B;
    const long long a = blockIdx.x / tiles;
    const int kx0 = (blockIdx.x % tiles) * B;
    const float2* col = P + a * M * M + kx0 + b;
    float2* s = lds + b * LS;
    auto gl = [&](int y) -> float2 { return col[(long long)y * M]; };
    auto gs = [&](int k, float2 v) { s[lds_pad(k)] = v; };
    Plan<M>::template run<-1>(t, s, tw, gl, gs);
    __syncthreads();
    float2* dst = spec + a * M * M;
    for (int e = threadIdx.x; e < B * M; e += B * T) {
        const int bb = e / M, ky = e % M, kx = kx0 + bb;
        float2 v = lds[bb * LS + lds_pad(ky)];
        if ((kx + ky) & 1) v = make_float2(-v.x, -v.y);
        dst[(long long)kx * M + ky] = v;
    }
}

template <int M>
static void rytov_cols_launch(Ctx& c, const float2* P, int A, float2* spec)
{
    constexpr int B = (256 / Plan<M>::T) < M ? (256 / Plan<M>::T) : M;  // columns per workgroup, <= M
    hipLaunchKernelGGL((k_rytov_cols<M, B>), dim3(A * (M / B)), dim3(B * Plan<M>::T), 0, c.stream, P, A,
                       c.tw(M), spec);
    CTDM_CHECK_LAUNCH("k_rytov_cols");
}

void rytov_spectrum_cols(Ctx& c, const float2* P, int A, int M, float2* spec)
{
    CTDM_REQUIRE(P && spec, "rytov_spectrum_cols: null pointer");
    if (A == 0) return;
    switch (M) {
    case 32: rytov_cols_launch<32>(c, P, A, spec); break;
    case 64: rytov_cols_launch<64>(c, P, A, spec); break;
    case 128: rytov_cols_launch<128>(c, P, A, spec); break;
    case 192: rytov_cols_launch<192>(c, P, A, spec); break;
    case 256: rytov_cols_launch<256>(c, P, A, spec); break;
    default: fail(CTDM_EUNSUPPORTED, "rytov: size has no kernel");
    }
}

template <int M>
static void rytov_launch(Ctx& c, const float* phi, const float* la, int A, float2* spec)
{
    constexpr int G = 256 / Plan<M>::T;
    constexpr int B = (256 / Plan<M>::T) < M ? (256 / Plan<M>::T) : M;  // columns per workgroup, <= M
    float2* P = (float2*)c.buf("ry.P", sizeof(float2) * (size_t)A * M * M);
    hipLaunchKernelGGL((k_rytov_rows<M, G>), dim3(grid1d((long long)A * M, G)), dim3(G * Plan<M>::T), 0, c.stream,
                       phi, la, A, c.tw(M), P);
    hipLaunchKernelGGL((k_rytov_cols<M, B>), dim3(A * (M / B)), dim3(B * Plan<M>::T), 0, c.stream, P, A,
                       c.tw(M), spec);
    CTDM_CHECK_LAUNCH("rytov");
}

void rytov_spectrum(Ctx& c, const float* unwrapped, const float* logamp, int A, int M, float2* spec)
{
    CTDM_REQUIRE(unwrapped && logamp && spec, "rytov_spectrum: null pointer");
    if (A == 0) return;
    switch (M) {
    case 32: rytov_launch<32>(c, unwrapped, logamp, A, spec); break;
    case 64: rytov_launch<64>(c, unwrapped, logamp, A, spec); break;
    case 128: rytov_launch<128>(c, unwrapped, logamp, A, spec); break;
    case 192: rytov_launch<192>(c, unwrapped, logamp, A, spec); break;
    case 256: rytov_launch<256>(c, unwrapped, logamp, A, spec); break;
    default: fail(CTDM_EUNSUPPORTED, "rytov: size has no kernel");
    }
}

// Reference-structured Rytov spectrum of one angle: cuCopyArray (Recon_kernel.cu:154-169) with the
// amplitude map, cuFFT2Dshift, cuFFT2D(-1).
__global__ void k_copy_array(float2* __restrict__ u, const float* __restrict__ phase, const float* __restrict__ amp,
                             int M)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= M || j >= M) return;
    const int idx = i + M * j;
    u[j + M * i] = make_float2(logf(amp[idx]), phase[idx]);
}

void rytov_spectrum_plain_one(Ctx& c, const float* phase, const float* amp, int M, float2* u_sp)
{
    dim3 blk(32, 8), grd((M + 31) / 32, (M + 7) / 8);
    hipLaunchKernelGGL(k_copy_array, grd, blk, 0, c.stream, u_sp, phase, amp, M);
    CTDM_CHECK_LAUNCH("k_copy_array");
    fftshift2d(c, u_sp, M, M, 1);
    fft_rows(c, u_sp, M, M, -1, 1.f, false);
    fft_cols(c, u_sp, M, M, M, 1, 0, -1, 1.f, false);
}

// ---------------------------------------------------------------- Ewald scatter / gather
// Owner-computes scatter: one thread per k-space pencil (Nx, Ny) walks the angles in order and
// updates its own pencil, so no atomics are needed and the summation order is fixed.
__global__ void k_scatter_owner(const float2* __restrict__ spec, const double* __restrict__ rows, int A,
                                int V, double fm02, double df, float2* __restrict__ F, float* __restrict__ hits)
{
    const long long VV = (long long)V * V;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= VV) return;
    const int Nx = (int)(p % V), Ny = (int)(p / V);
    const int ii = signed_freq(Nx, V), jj = signed_freq(Ny, V);
    for (int a = 0; a < A; ++a) {
        const Cap cp = ewald_cap(ii, jj, rows[3 * a], rows[3 * a + 1], rows[3 * a + 2], fm02, df, V);
        if (!cp.ok) continue;
        const float2 u = spec[(long long)a * VV + p];
        const long long v = p + VV * cp.nz;
        float2 f = F[v];
        f.x += -cp.w * u.y;
        f.y += cp.w * u.x;
        F[v] = f;
        hits[v] += 1.f;
    }
}

// fillEdwardSphere of one angle (Recon_kernel.cu:187-229): one thread per spectrum pixel, plain
// read-modify-write (the voxels of one angle are distinct).
__global__ void k_scatter_one(const float2* __restrict__ u_sp, double fx0, double fy0, double fz0, int V,
                              double fm02, double df, float2* __restrict__ F, float* __restrict__ hits)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= V || j >= V) return;
    const int ii = i - V / 2, jj = j - V / 2;
    const Cap cp = ewald_cap(ii, jj, fx0, fy0, fz0, fm02, df, V);
    if (!cp.ok) return;
    const int Nx = ii < 0 ? ii + V : ii, Ny = jj < 0 ? jj + V : jj;
    const float2 u = u_sp[Nx + (long long)V * Ny];
    const long long v = Nx + (long long)V * Ny + (long long)V * V * cp.nz;
    float2 f = F[v];
    f.x += -cp.w * u.y;
    f.y += cp.w * u.x;
    F[v] = f;
    hits[v] += 1.f;
}

// Adjoint of the scatter: spec[a][p] = conj(4 pi i f_z) F[v] on the cap, 0 elsewhere.
__global__ void k_gather(const float2* __restrict__ F, const double* __restrict__ rows, int A, int V,
                         double fm02, double df, float2* __restrict__ spec)
{
    const long long VV = (long long)V * V;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= VV * A) return;
    const int a = (int)(g / VV);
    const long long p = g % VV;
    const int ii = signed_freq((int)(p % V), V), jj = signed_freq((int)(p / V), V);
    const Cap cp = ewald_cap(ii, jj, rows[3 * a], rows[3 * a + 1], rows[3 * a + 2], fm02, df, V);
    float2 out = make_float2(0.f, 0.f);
    if (cp.ok) {
        const float2 f = F[p + VV * cp.nz];
        out = make_float2(cp.w * f.y, -cp.w * f.x);
    }
    spec[g] = out;
}

void ewald_scatter(Ctx& c, const float2* spec, const ctdm_angle* ang, int A, int M, int V,
                   const ctdm_optics& o, float2* F, float* hits)
{
    CTDM_REQUIRE(spec && F && hits, "ewald_scatter: null pointer");
    CTDM_REQUIRE(M == V && V >= 4 && V % 2 == 0, "ewald_scatter: spectrum side must equal the volume side");
    const EwaldGeom g = ewald_geom(o, V);
    const double* rows = upload_rows(c, "ew.rows", angle_rows(g, ang, A));
    const long long VV = (long long)V * V;
    hipLaunchKernelGGL(k_scatter_owner, dim3(grid1d(VV, 256)), dim3(256), 0, c.stream, spec, rows, A, V, g.fm02,
                       g.df, F, hits);
    CTDM_CHECK_LAUNCH("k_scatter_owner");
}

void ewald_gather(Ctx& c, const float2* F, const ctdm_angle* ang, int A, int M, int V, const ctdm_optics& o,
                  float2* spec)
{
    CTDM_REQUIRE(spec && F, "ewald_gather: null pointer");
    CTDM_REQUIRE(M == V && V >= 4 && V % 2 == 0, "ewald_gather: spectrum side must equal the volume side");
    const EwaldGeom g = ewald_geom(o, V);
    const double* rows = upload_rows(c, "ew.rows", angle_rows(g, ang, A));
    const long long n = (long long)V * V * A;
    hipLaunchKernelGGL(k_gather, dim3(grid1d(n, 256)), dim3(256), 0, c.stream, F, rows, A, V, g.fm02, g.df, spec);
    CTDM_CHECK_LAUNCH("k_gather");
}

// ---------------------------------------------------------------- normalisation (Initial_F_F2)
__global__ void k_normalize(float2* __restrict__ F, const float* __restrict__ C, float2* __restrict__ Fm, long long n,
                            float dx)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 f = F[i];
    const float c = C[i];
    if (c != 0.f) f = make_float2(f.x / c, f.y / c);
    f = make_float2(f.x / dx, f.y / dx);
    F[i] = f;
    if (Fm) Fm[i] = f;
}

void normalize_measured(Ctx& c, float2* F, const float* hits, float2* Fm, int V, double dx)
{
    CTDM_REQUIRE(F && hits && dx > 0, "normalize_measured: bad argument");
    const long long n = (long long)V * V * V;
    hipLaunchKernelGGL(k_normalize, dim3(grid1d(n, 256)), dim3(256), 0, c.stream, F, hits, Fm, n, (float)dx);
    CTDM_CHECK_LAUNCH("k_normalize");
}

// ---------------------------------------------------------------- 3-D FFT
static void check_vol(int V)
{
    if (!(V == 32 || V == 64 || V == 128 || V == 192 || V == 256 || V == 512))
        fail(CTDM_EUNSUPPORTED, "volume side " + std::to_string(V) + " has no kernel");
}

// x, y, z passes; `scale` folded into the z pass.
static void fft3d_passes(Ctx& c, float2* vol, int V, int dir, float scale)
{
    const long long VV = (long long)V * V;
    fft_rows(c, vol, V, VV, dir, 1.f, false);
    fft_cols(c, vol, V, V, V, V, VV, dir, 1.f, false);
    fft_cols(c, vol, V, V, VV, V, V, dir, scale, false);
}

void fft3d(Ctx& c, float2* vol, int V, int dir)
{
    CTDM_REQUIRE(vol, "fft3d: null pointer");
    CTDM_REQUIRE(dir == -1 || dir == 1, "fft3d: dir must be -1 or +1");
    check_vol(V);
    const double s = dir > 0 ? 1.0 / ((double)V * V * V) : 1.0;
    fft3d_passes(c, vol, V, dir, (float)s);
}

// ---------------------------------------------------------------- pointwise operators
struct Proj {
    float k2, inv_k2, nm, inv_nm2;
};

static Proj make_proj(const EwaldGeom& g)
{
    return Proj{(float)g.k2, (float)(1.0 / g.k2), (float)g.n_med, (float)(1.0 / (g.n_med * g.n_med))};
}

// calculate_n3D (Recon_kernel.cu:305-342): n = n_m sqrt(1 - f/k^2) with Re n >= n_m, Im n >= 0.
// The square root is evaluated in the cancellation-free form: with h = (|z| + |c|) / 2 the larger
// component is sqrt(h) and the smaller |d| / (2 sqrt(h)); it equals the reference's
// (sqrt((|z|+c)/2), sign(d) sqrt((|z|-c)/2)) in exact arithmetic. The hardware square root and
// reciprocal square root (v_sqrt_f32, v_rsq_f32, about 1 ulp) replace the correctly rounded library
// sequences: the x pass that applies this per voxel is bound by vector ALU issue, and the error
// stays four orders below the 1e-5 RI bar.
__device__ __forceinline__ float2 est_n(float2 f, const Proj& p)
{
    const float c = 1.f - f.x * p.inv_k2;
    const float d = -f.y * p.inv_k2;
    const float r = __builtin_amdgcn_sqrtf(c * c + d * d);
    const float h = 0.5f * (r + fabsf(c));
    const float big = __builtin_amdgcn_sqrtf(h);
    const float small = h > 0.f ? 0.5f * fabsf(d) * __builtin_amdgcn_rsqf(h) : 0.f;
    float a, b;
    if (c >= 0.f) {
        a = big;
        b = copysignf(small, d);
    } else {
        a = small;
        b = copysignf(big, d);
    }
    if (d == 0.f) b = 0.f;
    float nr = a * p.nm, ni = b * p.nm;
    if (nr < p.nm) nr = p.nm;
    if (ni < 0.f) ni = 0.f;
    return make_float2(nr, ni);
}

// Modify_F3D (Recon_kernel.cu:370-383) with the caller's n_m^2 (HilbertTransform.cu:1962):
// f = -k^2 ((n_r^2 - n_i^2)/n_m^2 - 1) + i (-k^2 2 n_r n_i / n_m^2), written without cancellation.
__device__ __forceinline__ float2 modify_f(float2 n, const Proj& p)
{
    const float re = ((n.x - p.nm) * (n.x + p.nm) - n.y * n.y) * p.inv_nm2;
    const float im = 2.f * n.x * n.y * p.inv_nm2;
    return make_float2(-p.k2 * re, -p.k2 * im);
}

// modify_f(est_n(scale * f)) in one step, as the fused x pass applies it: with (a, b) = sqrt(1 - f/k^2) and
// the positivity clamps a >= 1, b >= 0 (Re n >= n_m, Im n >= 0 without the n_m factor),
// f' = -k^2 ((a - 1)(a + 1) - b^2) - i k^2 2ab, the same value as the two operators in exact arithmetic.
__device__ __forceinline__ float2 project(float2 f, const Proj& p, float scale)
{
    const float ik = scale * p.inv_k2;
    const float c = 1.f - f.x * ik;
    const float d = -f.y * ik;
    const float r = __builtin_amdgcn_sqrtf(c * c + d * d);
    const float h = 0.5f * (r + fabsf(c));
    const float big = __builtin_amdgcn_sqrtf(h);
    const float small = h > 0.f ? 0.5f * fabsf(d) * __builtin_amdgcn_rsqf(h) : 0.f;
    float a = c >= 0.f ? big : small;
    float b = d > 0.f ? (c >= 0.f ? small : big) : 0.f;  // b < 0 (d < 0) and b = 0 (d = 0) both clamp to 0
    a = fmaxf(a, 1.f);
    return make_float2(-p.k2 * ((a - 1.f) * (a + 1.f) - b * b), -p.k2 * (2.f * a * b));
}

__global__ void k_est_n(const float2* __restrict__ f, float2* __restrict__ n, long long cnt, Proj p)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) n[i] = est_n(f[i], p);
}

__global__ void k_modify(const float2* __restrict__ n, float2* __restrict__ f, long long cnt, Proj p)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) f[i] = modify_f(n[i], p);
}

// check_F3D (Recon_kernel.cu:396-412) with the hit mask (default) or the reference's test.
__global__ void k_check(float2* __restrict__ F, const float2* __restrict__ Fm, const float* __restrict__ hits,
                        long long cnt, int refcompat)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const float2 m = Fm[i];
    const bool meas = refcompat ? (m.x != 0.f && m.y != 0.f) : (hits[i] > 0.f);
    if (meas) F[i] = m;
}

__global__ void k_scale(float2* __restrict__ F, long long cnt, float s)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) F[i] = make_float2(F[i].x * s, F[i].y * s);
}

__global__ void k_fill(float* __restrict__ p, long long cnt, float v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = v;
}

// Streaming copy, 16 B per lane: the bandwidth yardstick the volume passes are measured against and the
// known-size kernel that calibrates the memory-traffic counters.
__global__ __launch_bounds__(256) void k_copy_vol(const float4* __restrict__ s, float4* __restrict__ d, long long n4)
{
    const long long stride = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n4; i += stride) d[i] = s[i];
}

static void fill(Ctx& c, void* p, long long nfloats, float v)
{
    hipLaunchKernelGGL(k_fill, dim3(grid1d(nfloats, 256)), dim3(256), 0, c.stream, (float*)p, nfloats, v);
    CTDM_CHECK_LAUNCH("k_fill");
}

void positivity_plain(Ctx& c, float2* f, const float2* Fm, const float* hits, float2* n, int V, int iters,
                      const ctdm_optics& o, unsigned flags)
{
    CTDM_REQUIRE(f && Fm && n, "positivity: null pointer");
    const bool ref = flags & CTDM_DC_REFCOMPAT;
    CTDM_REQUIRE(hits || ref, "positivity: hits needed unless CTDM_DC_REFCOMPAT");
    check_vol(V);
    const EwaldGeom g = ewald_geom(o, V);
    const Proj p = make_proj(g);
    const long long cnt = (long long)V * V * V;
    const dim3 grd(grid1d(cnt, 256)), blk(256);
    const float s = (float)(1.0 / ((double)V * V * V));
    hipLaunchKernelGGL(k_est_n, grd, blk, 0, c.stream, f, n, cnt, p);
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_modify, grd, blk, 0, c.stream, n, f, cnt, p);
        fft3d_passes(c, f, V, -1, 1.f);
        hipLaunchKernelGGL(k_check, grd, blk, 0, c.stream, f, Fm, hits, cnt, (int)ref);
        fft3d_passes(c, f, V, +1, 1.f);
        hipLaunchKernelGGL(k_scale, grd, blk, 0, c.stream, f, cnt, s);
        hipLaunchKernelGGL(k_est_n, grd, blk, 0, c.stream, f, n, cnt, p);
    }
    CTDM_CHECK_LAUNCH("positivity_plain");
    if (!(flags & CTDM_POS_NO_SHIFT)) fftshift3d(c, n, V);
}

// ---------------------------------------------------------------- fused volume passes
// Compact data-consistency list: for pencil p = x + V*y, cnt[p] entries (z, value), stored at
// p*cap + k (z as int16 in idx, value in val).
struct DcList {
    int* cnt;
    short* idx;
    float2* val;
    int cap;
    // The fit's objective (both null otherwise): hit count of every list entry, written by the gridding
    // kernel, and one double per workgroup for the partial sums of the objective (gridding: its constant
    // part; z pass: the residual part before the step).
    short* hit = nullptr;
    double* part = nullptr;
};

// Sum of one double per thread over a workgroup of NT threads, in thread order (deterministic); every
// thread gets the total. `red` is NT doubles of LDS.
template <int NT>
__device__ __forceinline__ double workgroup_sum(double v, double* red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < NT; ++i) t += red[i];
        red[0] = t;
    }
    __syncthreads();
    return red[0];
}

// Pencil tile geometry of the volume passes. AXIS 0: rows (x), 1: y, 2: z. B pencils per workgroup.
template <int V, int AXIS, int B>
struct Tile {
    static constexpr int T = Plan<V>::T;
    static constexpr bool BFAST = AXIS != 0;
    int b, t;
    long long base;  // element offset of this thread's pencil
    long long stride;
    int px, py;      // pencil coordinates (the two fixed axes)
    // Pencils of B < 16 side by side cover less than a 128-byte line per row; neighbouring tiles then
    // share lines and are kept on one XCD (xcd_block).
    static constexpr bool REMAP = BFAST && B * 8 < 128;
    static __device__ int block() { return REMAP ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x; }
    // rev: tiles in reverse order (the volume swept from the last plane back), so that a pass can start
    // on the planes the previous pass wrote last
    __device__ Tile(int rev = 0)
    {
        const int tid = threadIdx.x;
        const int blk = rev ? (int)gridDim.x - 1 - block() : block();
        b = BFAST ? tid % B : tid / T;
        t = BFAST ? tid / B : tid % T;
        const long long VV = (long long)V * V;
        if (AXIS == 0) {  // pencil (y, z): rows
            const long long r = (long long)blk * B + b;
            px = (int)(r % V);
            py = (int)(r / V);
            base = r * V;
            stride = 1;
        } else if (AXIS == 1) {  // pencil (x, z)
            const int tiles = V / B;
            px = (blk % tiles) * B + b;
            py = blk / tiles;
            base = px + VV * py;
            stride = V;
        } else {  // pencil (x, y)
            const int tiles = V / B;
            px = (blk % tiles) * B + b;
            py = blk / tiles;
            base = px + (long long)V * py;
            stride = VV;
        }
    }
};

// X pass. MODE 0: P + forward (input real space); 1: inverse, scale, P, forward;
// 2: inverse, scale, est_n, octant shift -> out; 3: est_n (+ shift) -> out.
template <int V, int B, int MODE>
__global__ __launch_bounds__(B* Plan<V>::T) void k_vol_x(float2* __restrict__ W, float2* __restrict__ out,
                                                         const float2* __restrict__ tw, Proj p, float scale,
                                                         int shift, int rev)
{
    constexpr int T = Plan<V>::T;
    __shared__ float2 lds[B * LdsSize<V>::value];
    __shared__ float2 twl[LdsSize<V>::value];
    TwStager<V, B * T> tws;
    tws.fetch(tw);
    Tile<V, 0, B> tl(rev);
    float2* s = lds + tl.b * LdsSize<V>::value;
    float2* row = W + tl.base;
    // modes 1 and 2: the row's elements are loaded before the table's LDS copy waits for the table
    StageRegs<V, T, Plan<V>::R0> rin;
    if constexpr (MODE == 1 || MODE == 2) load_first_stage<V, T, Plan<V>::R0>(rin, row, 1, tl.t);
    __builtin_amdgcn_sched_barrier(0);
    tws.put(twl);
    const TwPad twp{twl};
    // destination row of the final n (octant shift)
    const int h = shift ? V / 2 : 0;
    const long long orow = ((long long)((tl.px + h) % V) + (long long)V * ((tl.py + h) % V)) * V;
    auto put_n = [&](int x, float2 v) { out[orow + (x + h) % V] = v; };
    if constexpr (MODE == 0) {
        auto gl = [&](int x) -> float2 { return project(row[x], p, 1.f); };
        auto gs = [&](int k, float2 v) { row[k] = v; };
        Plan<V>::template run<-1>(tl.t, s, twp, gl, gs);
    } else if constexpr (MODE == 3) {
        for (int x = tl.t; x < V; x += Plan<V>::T) put_n(x, est_n(row[x], p));
    } else {
        if constexpr (MODE == 1) {
            auto gs = [&](int k, float2 v) { s[lds_pad(k)] = project(v, p, scale); };
            Plan<V>::template run<+1>(tl.t, s, twp, rin, gs);
            __syncthreads();
            auto gl2 = [&](int k) -> float2 { return s[lds_pad(k)]; };
            auto gs2 = [&](int k, float2 v) { row[k] = v; };
            Plan<V>::template run<-1>(tl.t, s, twp, gl2, gs2);
        } else {
            auto gs = [&](int k, float2 v) { put_n(k, est_n(cscale(v, scale), p)); };
            Plan<V>::template run<+1>(tl.t, s, twp, rin, gs);
        }
    }
}

// Z pass with data consistency: forward z-FFT, replace measured voxels, inverse z-FFT.
// Load order at kernel entry: the twiddle table and the pencil's list count, then the pencil's elements
// (into the registers of the transform's first stage), then the list entries below the count (DC_PRE
// per thread; longer lists finish in a loop) and the table's LDS copy. Those wait for the count and the
// table only, so the data loads are in flight meanwhile, and no entry beyond the count is fetched (a
// pencil's list averages a third of its capacity).
// With tau != 1 a measured voxel moves only part of the way: F <- F + tau (m - F), the fit's
// hit-preconditioned gradient step (see fit_pgd); tau == 1 is the exact replacement.
// JT (the fit with its objective): before the step, each measured voxel adds h |F - m|^2 / 2 (h its
// hit count, m the averaged measurement) to the workgroup's partial sum, written to dc.part[block].
template <int V, int B, bool JT>
__global__ __launch_bounds__(B* Plan<V>::T) void k_vol_zdc(float2* __restrict__ W, const float2* __restrict__ tw,
                                                           DcList dc, float tau)
{
    constexpr int T = Plan<V>::T;
    constexpr int DC_PRE = 4;
    constexpr int LS = LdsStride<V, B>::value;
    __shared__ float2 lds[B * LS];
    __shared__ float2 twl[LdsSize<V>::value];
    __shared__ double red[JT ? B * T : 1];
    TwStager<V, B * T> tws;
    tws.fetch(tw);
    // list prefetch: lanes walk the entries of one pencil (contiguous in the list)
    const int pb = threadIdx.x / T, k0 = threadIdx.x % T;
    const int blk = Tile<V, 2, B>::block();
    const long long dpencil = (long long)(blk % (V / B)) * B + pb + (long long)V * (blk / (V / B));
    const int n = min(dc.cnt[dpencil], dc.cap);  // a count never exceeds the list capacity
    Tile<V, 2, B> tl;
    float2* s = lds + tl.b * LS;
    float2* pen = W + tl.base;
    StageRegs<V, T, Plan<V>::R0> x;
    load_first_stage<V, T, Plan<V>::R0>(x, pen, tl.stride, tl.t);
    __builtin_amdgcn_sched_barrier(0);  // keeps the data loads ahead of the waits for the count and the table
    tws.put(twl);
    const TwPad twp{twl};
    // Branch-free: lanes past the count read the pencil's first entry (a line the others fetch anyway),
    // and the entries are used only below the count.
    int pz[DC_PRE], ph[DC_PRE];
    float2 pv[DC_PRE];
#pragma unroll
    for (int r = 0; r < DC_PRE; ++r) {
        const int k = k0 + r * T;
        const long long e = dpencil * dc.cap + (k < n ? k : 0);
        pz[r] = dc.idx[e];
        pv[r] = dc.val[e];
        if constexpr (JT) ph[r] = dc.hit[e];
    }
    double jacc = 0.0;
    auto gs = [&](int k, float2 v) { s[lds_pad(k)] = v; };
    Plan<V>::template run<-1>(tl.t, s, twp, x, gs);
    __syncthreads();
    {
        float2* sp = lds + pb * LS;
        const bool relax = tau != 1.f;
        auto put = [&](int z, float2 m, int h) {  // each z appears once in a pencil's list
            if (relax || JT) {
                const float2 F = sp[lds_pad(z)];
                if constexpr (JT) {
                    const double dr = (double)F.x - m.x, di = (double)F.y - m.y;
                    jacc += 0.5 * h * (dr * dr + di * di);
                }
                if (relax) m = make_float2(F.x + tau * (m.x - F.x), F.y + tau * (m.y - F.y));
            }
            sp[lds_pad(z)] = m;
        };
#pragma unroll
        for (int r = 0; r < DC_PRE; ++r) {
            const int k = k0 + r * T;
            const int z = pz[r];
            if (k < n && z >= 0 && z < V) put(z, pv[r], JT ? ph[r] : 0);
        }
        for (int k = k0 + DC_PRE * T; k < n; k += T) {
            const long long e = dpencil * dc.cap + k;
            const int z = dc.idx[e];
            if (z >= 0 && z < V) put(z, dc.val[e], JT ? dc.hit[e] : 0);
        }
    }
    if constexpr (JT) {
        const double tot = workgroup_sum<B * T>(jacc, red);  // its barriers also order the LDS updates
        if (threadIdx.x == 0) dc.part[blockIdx.x] = tot;
    }
    __syncthreads();
    auto gl2 = [&](int k) -> float2 { return s[lds_pad(k)]; };
    auto gs2 = [&](int z, float2 v) { pen[(long long)z * tl.stride] = v; };
    Plan<V>::template run<+1>(tl.t, s, twp, gl2, gs2);
}

// Y pass (plain FFT along y).
template <int V, int B, int DIR, int PV>
__global__ __launch_bounds__(B* Plan<V>::T) void k_vol_y(float2* __restrict__ W, const float2* __restrict__ tw,
                                                         int rev)
{
    using PL = std::conditional_t<PV == 0, Plan<V>, PlanAlt<V>>;
    constexpr int T = PL::T;
    constexpr int LS = LdsStride<V, B>::value;
    __shared__ float2 lds[B * LS];
    __shared__ float2 twl[LdsSize<V>::value];
    // twiddle table, then the pencil's elements, then the table's LDS copy: the data loads are issued
    // before the first wait for the table (the first stage's barrier publishes it)
    TwStager<V, B * T> tws;
    tws.fetch(tw);
    Tile<V, 1, B> tl(rev);
    float2* s = lds + tl.b * LS;
    float2* pen = W + tl.base;
    StageRegs<V, T, PL::R0> x;
    load_first_stage<V, T, PL::R0>(x, pen, tl.stride, tl.t);
    __builtin_amdgcn_sched_barrier(0);
    tws.put(twl);
    const TwPad twp{twl};
    auto gs = [&](int k, float2 v) { pen[(long long)k * tl.stride] = v; };
    PL::template run<DIR>(tl.t, s, twp, x, gs);
}

// Compact list from a dense measured spectrum (positivity API): one thread per pencil.
__global__ void k_dc_build(const float2* __restrict__ Fm, const float* __restrict__ hits, int V, int refcompat,
                           DcList dc)
{
    const long long VV = (long long)V * V;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= VV) return;
    int n = 0;
    for (int z = 0; z < V; ++z) {
        const long long v = p + VV * z;
        const float2 m = Fm[v];
        const bool meas = refcompat ? (m.x != 0.f && m.y != 0.f) : (hits[v] > 0.f);
        if (meas) {
            dc.idx[p * dc.cap + n] = (short)z;
            dc.val[p * dc.cap + n] = m;
            ++n;
        }
    }
    dc.cnt[p] = n;
}

// Gridding pass: per k-space pencil (x, y) accumulate every angle's sample in angle order, average
// by hit count, scale by 1/dx (Initial_F_F2), emit the compact list, then (VOLUME) inverse z-FFT
// into W.
// The pencil's T threads evaluate the caps of interleaved angles into an LDS list (chunks of
// GRID_CHUNK angles); thread t then owns the z indices t + T j and adds the list entries that land
// there in angle order, so the sums are deterministic. A pencil's list holds thread t's entries
// (ascending z) after those of threads below t; the data-consistency step only needs each measured z
// once, in any order.
template <int V, int B, bool VOLUME>
__global__ __launch_bounds__(B* Plan<V>::T) void k_vol_grid(const float2* __restrict__ spec,
                                                            const double* __restrict__ rows, int A,
                                                            double fm02, double df, float inv_dx,
                                                            int refcompat, const float2* __restrict__ tw,
                                                            float2* __restrict__ W, DcList dc)
{
    constexpr int T = Plan<V>::T;
    constexpr int APT = (GRID_CHUNK + T - 1) / T;  // angles per thread in one chunk
    constexpr int ZR = V / T;                        // z indices owned by one thread
    constexpr int HW = (ZR + 1) / 2;                 // 16-bit hit counters, two per register
    constexpr int LS = LdsStride<V, B>::value;
    __shared__ float2 lds[B * LS];
    __shared__ short lnz[B][GRID_CHUNK];
    __shared__ float2 lval[B][GRID_CHUNK + 1];
    __shared__ unsigned lmask[B][T];  // owner masks of the staged chunk
    __shared__ double red[VOLUME ? 1 : B * T];  // the fit's objective constant (list only)
    Tile<V, 2, B> tl;
    const long long VV = (long long)V * V;
    const long long pencil = tl.base;  // x + V*y
    // This thread's share of a chunk of angles: the angles a0 + t + T i. fetch() issues the loads
    // (cap centres first, then spectrum samples) into registers, stage() evaluates the caps into the
    // LDS list and marks each entry's owner. The first chunk is fetched before the twiddle staging and
    // the zeroing, every later one during the previous chunk's accumulation, so the spectrum's HBM
    // latency overlaps other work.
    float2 u[APT];
    double fx0[APT], fy0[APT], fz0[APT];
    auto fetch = [&](int a0) {
        const int na = min(GRID_CHUNK, A - a0);
#pragma unroll
        for (int i = 0; i < APT; ++i) {
            const int k = tl.t + i * T;
            const int a = a0 + (k < na ? k : 0);
            fx0[i] = rows[3 * a];
            fy0[i] = rows[3 * a + 1];
            fz0[i] = rows[3 * a + 2];
        }
#pragma unroll
        for (int i = 0; i < APT; ++i) {
            const int k = tl.t + i * T;
            const int a = a0 + (k < na ? k : 0);
            u[i] = spec[(long long)a * VV + pencil];
        }
    };
    const int ii = signed_freq(tl.px, V), jj = signed_freq(tl.py, V);
    auto stage = [&](int a0) {
        const int na = min(GRID_CHUNK, A - a0);
#pragma unroll
        for (int i = 0; i < APT; ++i) {
            const int k = tl.t + i * T;
            if (k < na) {
                const Cap cp = ewald_cap(ii, jj, fx0[i], fy0[i], fz0[i], fm02, df, V);
                short nz = -1;
                float2 val = make_float2(0.f, 0.f);
                if (cp.ok) {
                    nz = (short)cp.nz;
                    val = make_float2(-cp.w * u[i].y, cp.w * u[i].x);
                }
                lnz[tl.b][k] = nz;
                lval[tl.b][k] = val;
                if (nz >= 0) mark_owner<T>(lmask[tl.b], nz, k);
            }
        }
    };
    if (A > 0) fetch(0);
    float2* s = lds + tl.b * LS;
    for (int z = tl.t; z < V; z += T) s[lds_pad(z)] = make_float2(0.f, 0.f);  // owned z indices
    // Hit counts of the owned z indices t + T j, in registers: counter j is bits 16 (j & 1) of hc[j / 2].
    unsigned hc[HW];
#pragma unroll
    for (int w = 0; w < HW; ++w) hc[w] = 0;
    double sq = 0.0;  // list only: sum of |sample|^2 over this thread's entries
    lmask[tl.b][tl.t] = 0;
    __syncthreads();  // cleared masks before any mark (the loads stay in flight across the barrier)
    for (int a0 = 0; a0 < A; a0 += GRID_CHUNK) {
        if (a0 > 0) __syncthreads();  // every owner is done with the previous chunk and cleared its mask
        stage(a0);
        if (a0 + GRID_CHUNK < A) fetch(a0 + GRID_CHUNK);  // in flight during this chunk's accumulation
        __syncthreads();
        unsigned mine = lmask[tl.b][tl.t];  // this thread's entries, in angle order
        lmask[tl.b][tl.t] = 0;
        while (mine) {
            const int k = __ffs(mine) - 1;
            mine &= mine - 1;
            const int nz = lnz[tl.b][k];
            float2& acc = s[lds_pad(nz)];
            const float2 v = lval[tl.b][k];
            acc.x += v.x;
            acc.y += v.y;
            if constexpr (!VOLUME) sq += (double)v.x * v.x + (double)v.y * v.y;
            const int j = nz / T;
#pragma unroll
            for (int w = 0; w < HW; ++w)
                if (w == (j >> 1)) hc[w] += 1u << (16 * (j & 1));
        }
    }
    auto hits = [&](int j) -> int { return (int)((hc[j >> 1] >> (16 * (j & 1))) & 0xffffu); };
    // normalise the owned z indices (each touched only by its owner so far) and count list entries:
    // the mean over the h hits times 1/dx, as one scale inv_dx / h with the hardware reciprocal
    // (v_rcp_f32, about 1 ulp) in place of two correctly rounded divisions per voxel
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < ZR; ++j) {
        const int z = tl.t + j * T;
        const int h = hits(j);
        float2 f = s[lds_pad(z)];
        const float sc = h ? inv_dx * __builtin_amdgcn_rcpf((float)h) : inv_dx;
        f = make_float2(f.x * sc, f.y * sc);
        s[lds_pad(z)] = f;
        if (h && !(refcompat && !(f.x != 0.f && f.y != 0.f))) ++cnt;
    }
    lmask[tl.b][tl.t] = cnt;  // the owner's own (cleared) mask word holds its list count from here on
    __syncthreads();
    int off = 0;
    for (int k = 0; k < tl.t; ++k) off += (int)lmask[tl.b][k];
    double hm2 = 0.0;  // list only: sum of h |m|^2 over this thread's measured voxels
#pragma unroll
    for (int j = 0; j < ZR; ++j) {
        const int z = tl.t + j * T;
        const int h = hits(j);
        const float2 f = s[lds_pad(z)];
        if (h && !(refcompat && !(f.x != 0.f && f.y != 0.f))) {
            dc.idx[pencil * dc.cap + off] = (short)z;
            dc.val[pencil * dc.cap + off] = f;
            if constexpr (!VOLUME) {
                if (dc.hit) dc.hit[pencil * dc.cap + off] = (short)h;
                hm2 += h * ((double)f.x * f.x + (double)f.y * f.y);
            }
            ++off;
        }
    }
    if (tl.t == T - 1) dc.cnt[pencil] = off;
    if constexpr (!VOLUME) {  // list only (the fit's measurement list; W is not touched)
        if (dc.part) {
            // objective constant: sum over voxels of (sum |m|^2 - h |m-bar|^2) / 2, m = sample / dx
            const double idx2 = (double)inv_dx * inv_dx;
            const double tot = workgroup_sum<B * T>(0.5 * (sq * idx2 - hm2), red);
            if (threadIdx.x == 0) dc.part[blockIdx.x] = tot;
        }
        return;
    }
    float2* pen = W + tl.base;
    auto gl = [&](int k) -> float2 { return s[lds_pad(k)]; };
    auto gs = [&](int z, float2 v) { pen[VV * z] = v; };
    Plan<V>::template run<+1>(tl.t, s, tw, gl, gs);  // twiddles from the (cached) global table
}

static DcList list_alloc(Ctx& c, const std::string& prefix, int V, int cap)
{
    const size_t VV = (size_t)V * V;
    DcList d;
    d.cap = cap;
    d.cnt = (int*)c.buf(prefix + ".cnt", sizeof(int) * VV);
    d.idx = (short*)c.buf(prefix + ".idx", sizeof(short) * VV * cap);
    d.val = (float2*)c.buf(prefix + ".val", sizeof(float2) * VV * cap);
    return d;
}

static DcList dc_alloc(Ctx& c, int V, int cap)
{
    if (c.lane == 0) {  // the list the pass-timing helpers reuse is lane 0's
        c.dc_V = V;
        c.dc_cap = cap;
    }
    return list_alloc(c, "dc", V, cap);
}

// Launchers of the fused passes for one pencils-per-workgroup choice B.
// Launches a volume pass. When the context holds a pair of timing events they are attached to this
// dispatch (recorded as it starts and as it ends, no packets in between) and consumed.
template <class K, class... Args>
static void vlaunch(Ctx& c, K kernel, dim3 g, dim3 b, Args... args)
{
    if (c.tev[0]) {
        hipExtLaunchKernelGGL(kernel, g, b, 0, c.stream, c.tev[0], c.tev[1], 0, args...);
        c.tev[0] = c.tev[1] = nullptr;
    } else {
        hipLaunchKernelGGL(kernel, g, b, 0, c.stream, args...);
    }
}

template <int V, int B>
struct VolOps {
    static constexpr int NT = B * Plan<V>::T;
    static dim3 grid() { return dim3((unsigned)((long long)V * V / B)); }

    static void x(Ctx& c, int mode, float2* W, float2* out, const Proj& p, float scale, bool shift, int rev = 0)
    {
        const float2* tw = c.tw(V);
        const int sh = shift;
        switch (mode) {
        case 0: vlaunch(c, &k_vol_x<V, B, 0>, grid(), dim3(NT), W, out, tw, p, scale, sh, rev); break;
        case 1: vlaunch(c, &k_vol_x<V, B, 1>, grid(), dim3(NT), W, out, tw, p, scale, sh, rev); break;
        case 2: vlaunch(c, &k_vol_x<V, B, 2>, grid(), dim3(NT), W, out, tw, p, scale, sh, rev); break;
        default: vlaunch(c, &k_vol_x<V, B, 3>, grid(), dim3(NT), W, out, tw, p, scale, sh, rev); break;
        }
        CTDM_CHECK_LAUNCH("k_vol_x");
    }
    static void y(Ctx& c, float2* W, int dir, int rev = 0, int plan = 0)
    {
        const float2* tw = c.tw(V);
        if (plan == 0) {
            if (dir < 0)
                vlaunch(c, &k_vol_y<V, B, -1, 0>, grid(), dim3(NT), W, tw, rev);
            else
                vlaunch(c, &k_vol_y<V, B, +1, 0>, grid(), dim3(NT), W, tw, rev);
        } else {
            if (dir < 0)
                vlaunch(c, &k_vol_y<V, B, -1, 1>, grid(), dim3(NT), W, tw, rev);
            else
                vlaunch(c, &k_vol_y<V, B, +1, 1>, grid(), dim3(NT), W, tw, rev);
        }
        CTDM_CHECK_LAUNCH("k_vol_y");
    }
    static void zdc(Ctx& c, float2* W, const DcList& dc, float tau)
    {
        const float2* tw = c.tw(V);
        if (dc.part && dc.hit)
            vlaunch(c, &k_vol_zdc<V, B, true>, grid(), dim3(NT), W, tw, dc, tau);
        else
            vlaunch(c, &k_vol_zdc<V, B, false>, grid(), dim3(NT), W, tw, dc, tau);
        CTDM_CHECK_LAUNCH("k_vol_zdc");
    }
    // W == nullptr: the list only.
    static void grid_pass(Ctx& c, const float2* spec, const double* rows, int A, const EwaldGeom& g, bool ref,
                          float2* W, const DcList& dc)
    {
        if (W)
            hipLaunchKernelGGL((k_vol_grid<V, B, true>), grid(), dim3(NT), 0, c.stream, spec, rows, A, g.fm02,
                               g.df, (float)(1.0 / g.dx), (int)ref, c.tw(V), W, dc);
        else
            hipLaunchKernelGGL((k_vol_grid<V, B, false>), grid(), dim3(NT), 0, c.stream, spec, rows, A, g.fm02,
                               g.df, (float)(1.0 / g.dx), (int)ref, c.tw(V), W, dc);
        CTDM_CHECK_LAUNCH("k_vol_grid");
    }
};

// Runtime choice of B per pass kind (autotuned; 16 by default).
enum PassKind { PASS_X = 0, PASS_Y = 1, PASS_ZDC = 2, PASS_GRID = 3, NPASS = 4, PASS_COPY = 4 };

template <int V>
struct Vol {
    static int pick(Ctx& c, PassKind k) { return c.tuning(V).pencils[k] == 8 ? 8 : 16; }
    // Sweep direction of the x and y passes (tuned): both walk the z planes in tile order, so running
    // one of them backwards starts each pass on the planes the previous one wrote last.
    static void x(Ctx& c, int mode, float2* W, float2* out, const Proj& p, float scale, bool shift)
    {
        const int rev = c.tuning(V).rev[0];
        if (pick(c, PASS_X) == 8) VolOps<V, 8>::x(c, mode, W, out, p, scale, shift, rev);
        else VolOps<V, 16>::x(c, mode, W, out, p, scale, shift, rev);
    }
    static void y(Ctx& c, float2* W, int dir)
    {
        const int rev = c.tuning(V).rev[1], plan = c.tuning(V).yplan;
        if (pick(c, PASS_Y) == 8) VolOps<V, 8>::y(c, W, dir, rev, plan);
        else VolOps<V, 16>::y(c, W, dir, rev, plan);
    }
    static void zdc(Ctx& c, float2* W, const DcList& dc, float tau = 1.f)
    {
        if (pick(c, PASS_ZDC) == 8) VolOps<V, 8>::zdc(c, W, dc, tau);
        else VolOps<V, 16>::zdc(c, W, dc, tau);
    }
    static void grid_pass(Ctx& c, const float2* spec, const double* rows, int A, const EwaldGeom& g, bool ref,
                          float2* W, const DcList& dc)
    {
        if (pick(c, PASS_GRID) == 8) VolOps<V, 8>::grid_pass(c, spec, rows, A, g, ref, W, dc);
        else VolOps<V, 16>::grid_pass(c, spec, rows, A, g, ref, W, dc);
    }
    // Loop after the initial inverse transform of the measured spectrum is in W up to the x pass:
    // `first_x_mode` 0 if W holds real-space f, 1 if W still needs its inverse x-FFT. The z step is
    // the data-consistency replacement (tau = 1) or the fit's relaxed step F <- F + tau (m - F).
    // `jpart` (the fit with its objective): iteration it's z pass writes its workgroup partial sums to
    // jpart + it * jstride.
    static void loop(Ctx& c, float2* W, float2* n_out, const DcList& dc, const Proj& p, int iters, bool shift,
                     int first_x_mode, float tau = 1.f, double* jpart = nullptr, int jstride = 0)
    {
        const float s = (float)(1.0 / ((double)V * V * V));
        if (iters == 0) {
            if (first_x_mode == 0)
                x(c, 3, W, n_out, p, 1.f, shift);
            else
                x(c, 2, W, n_out, p, s, shift);
            return;
        }
        x(c, first_x_mode, W, n_out, p, s, shift);
        for (int it = 0; it < iters; ++it) {
            y(c, W, -1);
            DcList d = dc;
            d.part = jpart ? jpart + (size_t)it * jstride : nullptr;
            zdc(c, W, d, tau);
            y(c, W, +1);
            x(c, it + 1 < iters ? 1 : 2, W, n_out, p, s, shift);
        }
    }
};

template <int V>
static void positivity_fused_v(Ctx& c, float2* f, const float2* Fm, const float* hits, float2* n, int iters,
                               const Proj& p, unsigned flags)
{
    const bool ref = flags & CTDM_DC_REFCOMPAT;
    const DcList dc = dc_alloc(c, V, V);
    hipLaunchKernelGGL(k_dc_build, dim3(grid1d((long long)V * V, 256)), dim3(256), 0, c.stream, Fm, hits, V,
                       (int)ref, dc);
    CTDM_CHECK_LAUNCH("k_dc_build");
    Vol<V>::loop(c, f, n, dc, p, iters, !(flags & CTDM_POS_NO_SHIFT), 0);
}

void positivity_fused(Ctx& c, float2* f, const float2* Fm, const float* hits, float2* n, int V, int iters,
                      const ctdm_optics& o, unsigned flags)
{
    CTDM_REQUIRE(f && Fm && n && f != n, "positivity: bad pointers");
    const bool ref = flags & CTDM_DC_REFCOMPAT;
    CTDM_REQUIRE(hits || ref, "positivity: hits needed unless CTDM_DC_REFCOMPAT");
    CTDM_REQUIRE(iters >= 0, "positivity: negative iteration count");
    const Proj p = make_proj(ewald_geom(o, V));
    switch (V) {
    case 32: positivity_fused_v<32>(c, f, Fm, hits, n, iters, p, flags); break;
    case 64: positivity_fused_v<64>(c, f, Fm, hits, n, iters, p, flags); break;
    case 128: positivity_fused_v<128>(c, f, Fm, hits, n, iters, p, flags); break;
    case 192: positivity_fused_v<192>(c, f, Fm, hits, n, iters, p, flags); break;
    case 256: positivity_fused_v<256>(c, f, Fm, hits, n, iters, p, flags); break;
    default: fail(CTDM_EUNSUPPORTED, "fused positivity: volume side has no kernel");
    }
}

template <int V>
static void volume_batched_v(Ctx& c, const float2* spec, const double* rows, int A, const EwaldGeom& g, int iters,
                             unsigned flags, float2* n_out)
{
    const DcList dc = dc_alloc(c, V, std::max(A, 1));
    float2* W = (float2*)c.buf("vol.W", sizeof(float2) * (size_t)V * V * V);
    Vol<V>::grid_pass(c, spec, rows, A, g, flags & CTDM_DC_REFCOMPAT, W, dc);
    Vol<V>::y(c, W, +1);
    Vol<V>::loop(c, W, n_out, dc, make_proj(g), iters, !(flags & CTDM_POS_NO_SHIFT), 1);
}

void volume_batched(Ctx& c, const float2* spec, const double* rows, int A, int V, const ctdm_optics& o,
                    int iters, unsigned flags, float2* n_out)
{
    CTDM_REQUIRE(spec && n_out && (rows || A == 0), "volume: null pointer");
    CTDM_REQUIRE(A >= 0 && A < 32768 && iters >= 0, "volume: bad sizes");
    const EwaldGeom g = ewald_geom(o, V);
    switch (V) {
    case 32: volume_batched_v<32>(c, spec, rows, A, g, iters, flags, n_out); break;
    case 64: volume_batched_v<64>(c, spec, rows, A, g, iters, flags, n_out); break;
    case 128: volume_batched_v<128>(c, spec, rows, A, g, iters, flags, n_out); break;
    case 192: volume_batched_v<192>(c, spec, rows, A, g, iters, flags, n_out); break;
    case 256: volume_batched_v<256>(c, spec, rows, A, g, iters, flags, n_out); break;
    default: fail(CTDM_EUNSUPPORTED, "volume side has no kernel");
    }
}

// Projected-gradient fit. With F = FFT3(f), m_a(p) = 4 pi i f_z spec_a(p) / dx (the sample the reference
// scatters, Recon_kernel.cu:224-225, with the 1/dx of Initial_F_F2, :269-270) and G the unit-weight
// Ewald gather (F -> F[v_a(p)], adjoint G^T = the scatter), J(f) = 1/2 sum_a |G_a F - m_a|^2 has the
// k-space gradient R = G^T (G F - m). G selects one voxel per (angle, pencil), so G^T G = H, the
// diagonal of hit counts, and R = H F - G^T m. The hit-preconditioned step F <- F - tau H^-1 R is
// therefore F <- F + tau (m-bar - F) on the measured voxels, with m-bar = H^-1 G^T m the averaged
// measurement. m-bar is formed once by the gridding kernel (list only); each iteration then recomputes
// the 3-D FFT and runs the relaxed data-consistency pass, then positivity (est_n) in real space.
// tau = 1 is the reference's loop (check_F3D replacement).
// out[r] = base[0] (if given) + the sum of part[r * n .. r * n + n), one workgroup per row r, in a fixed
// order (deterministic).
__global__ __launch_bounds__(256) void k_sum_rows(const double* __restrict__ part, int n, const double* base,
                                                  double* __restrict__ out)
{
    __shared__ double red[256];
    const double* row = part + (size_t)blockIdx.x * n;
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) t += row[i];
    const double tot = workgroup_sum<256>(t, red);
    if (threadIdx.x == 0) out[blockIdx.x] = tot + (base ? base[0] : 0.0);
}

// objective (host, may be null): J before each step, J_k = 1/2 sum_a sum_p |FFT3(P f_k)[v_a(p)] - m_a(p)|^2
// with P the positivity projection. Per voxel the samples' residuals sum to h |F - m-bar|^2 + (sum |m|^2 -
// h |m-bar|^2), so the z pass accumulates the first term and the gridding kernel the constant second one.
template <int V>
static void fit_pgd_v(Ctx& c, float2* f, const float2* spec, const double* rows, int A, const EwaldGeom& g,
                      int iters, float tau, float2* n, unsigned flags, double* objective)
{
    DcList ml = list_alloc(c, "pgd", V, std::max(A, 1));
    double* jpart = nullptr;
    double* jdev = nullptr;
    int jstride = 0;
    if (objective && iters > 0) {
        const int gblk = V * V / Vol<V>::pick(c, PASS_GRID), zblk = V * V / Vol<V>::pick(c, PASS_ZDC);
        jstride = zblk;
        ml.hit = (short*)c.buf("pgd.hit", sizeof(short) * (size_t)V * V * ml.cap);
        ml.part = (double*)c.buf("pgd.cpart", sizeof(double) * gblk);
        jpart = (double*)c.buf("pgd.jpart", sizeof(double) * (size_t)iters * zblk);
        jdev = (double*)c.buf("pgd.J", sizeof(double) * (iters + 1));
        Vol<V>::grid_pass(c, spec, rows, A, g, false, nullptr, ml);
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, c.stream, ml.part, gblk, (const double*)nullptr, jdev);
        CTDM_CHECK_LAUNCH("k_sum_rows");
        ml.part = nullptr;
    } else {
        Vol<V>::grid_pass(c, spec, rows, A, g, false, nullptr, ml);  // m-bar = H^-1 G^T m per measured voxel
    }
    Vol<V>::loop(c, f, n, ml, make_proj(g), iters, !(flags & CTDM_POS_NO_SHIFT), 0, tau, jpart, jstride);
    if (jpart) {
        hipLaunchKernelGGL(k_sum_rows, dim3(iters), dim3(256), 0, c.stream, jpart, jstride, (const double*)jdev,
                           jdev + 1);
        CTDM_CHECK_LAUNCH("k_sum_rows");
        CTDM_HIP(hipMemcpyAsync(objective, jdev + 1, sizeof(double) * iters, hipMemcpyDeviceToHost, c.stream));
        CTDM_HIP(hipStreamSynchronize(c.stream));
    }
}

void fit_pgd(Ctx& c, float2* f, const float2* spec, const ctdm_angle* ang, int A, int V, const ctdm_optics& o,
             int iters, double step, float2* n, unsigned flags, double* objective)
{
    CTDM_REQUIRE(f && spec && n && f != n, "fit_pgd: bad pointers");
    CTDM_REQUIRE(A > 0 && A < 32768 && iters >= 0, "fit_pgd: bad sizes");
    CTDM_REQUIRE(step > 0 && step <= 2, "fit_pgd: step must lie in (0, 2]");
    const EwaldGeom g = ewald_geom(o, V);
    const double* rows = upload_rows(c, "pgd.rows", angle_rows(g, ang, A));
    switch (V) {
    case 32: fit_pgd_v<32>(c, f, spec, rows, A, g, iters, (float)step, n, flags, objective); break;
    case 64: fit_pgd_v<64>(c, f, spec, rows, A, g, iters, (float)step, n, flags, objective); break;
    case 128: fit_pgd_v<128>(c, f, spec, rows, A, g, iters, (float)step, n, flags, objective); break;
    case 192: fit_pgd_v<192>(c, f, spec, rows, A, g, iters, (float)step, n, flags, objective); break;
    case 256: fit_pgd_v<256>(c, f, spec, rows, A, g, iters, (float)step, n, flags, objective); break;
    default: fail(CTDM_EUNSUPPORTED, "fit_pgd: volume side has no kernel");
    }
}

// Scatter of one angle's spectrum, reference-structured (used by the plain pipeline).
void scatter_plain_one(Ctx& c, const float2* u_sp, const AngleRow& r, int V, const EwaldGeom& g, float2* F,
                       float* hits)
{
    dim3 blk(32, 8), grd((V + 31) / 32, (V + 7) / 8);
    hipLaunchKernelGGL(k_scatter_one, grd, blk, 0, c.stream, u_sp, r.fx0, r.fy0, r.fz0, V, g.fm02, g.df, F, hits);
    CTDM_CHECK_LAUNCH("k_scatter_one");
}

void zero_volume(Ctx& c, float2* F, float* hits, int V)
{
    const long long n = (long long)V * V * V;
    fill(c, F, 2 * n, 0.f);
    if (hits) fill(c, hits, n, 0.f);
}

// ---------------------------------------------------------------- pass timing (autotuner, bench)
// Average duration of one launch of pass `kind` with B pencils per workgroup, timed with hipEvents on
// the context stream over `reps` launches. Runs on the context's volume buffer and data-consistency
// list (left by the last batched reconstruction of side V, or zero-initialised here). Returns the
// algorithmic bytes of one launch (logical inputs read once + outputs written once).
template <int V>
static double time_pass_v(Ctx& c, int kind, int B, int reps, double* bytes)
{
    const size_t vol = (size_t)V * V * V, VV = (size_t)V * V;
    const std::string wname = "vol.W";
    const bool have = c.dc_V == V && c.scratch.count(wname) && c.scratch[wname].bytes >= sizeof(float2) * vol;
    const int cap = have ? std::max(1, c.dc_cap) : 1;
    float2* W = (float2*)c.buf(wname, sizeof(float2) * vol);
    const DcList dc = dc_alloc(c, V, cap);
    if (!have) {
        fill(c, W, 2 * (long long)vol, 1.f);
        fill(c, dc.cnt, (long long)VV, 0.f);  // all-zero bit pattern == integer 0
    }
    float2* nbuf = (float2*)c.buf("tp.n", sizeof(float2) * vol);
    const Proj p = make_proj(ewald_geom(ctdm_optics{532e-9, 1.337, 0.4e-6}, V));
    // bytes of one launch
    const double vb = 2.0 * sizeof(float2) * vol;
    double extra = 0;
    if (kind == PASS_ZDC) {
        std::vector<int> h(VV);
        CTDM_HIP(hipMemcpy(h.data(), dc.cnt, sizeof(int) * VV, hipMemcpyDeviceToHost));
        double n = 0;
        for (int v : h) n += v;
        extra = sizeof(int) * (double)VV + n * (sizeof(float2) + sizeof(short));
    }
    *bytes = vb + extra;
    const long long n4 = (long long)vol / 2;  // float4 = two complex values
    const int copy_grid = (int)std::min<long long>(grid1d(n4, 256), 8192);
    auto launch = [&](int r) {
        switch (kind) {
        case PASS_COPY:
            hipLaunchKernelGGL(k_copy_vol, dim3(copy_grid), dim3(256), 0, c.stream, (const float4*)W, (float4*)nbuf, n4);
            CTDM_CHECK_LAUNCH("k_copy_vol");
            break;
        case PASS_X: B == 8 ? VolOps<V, 8>::x(c, 1, W, nbuf, p, 1.f / (float)vol, true)
                            : VolOps<V, 16>::x(c, 1, W, nbuf, p, 1.f / (float)vol, true); break;
        case PASS_Y: B == 8 ? VolOps<V, 8>::y(c, W, (r & 1) ? 1 : -1) : VolOps<V, 16>::y(c, W, (r & 1) ? 1 : -1); break;
        case PASS_ZDC: B == 8 ? VolOps<V, 8>::zdc(c, W, dc, 1.f) : VolOps<V, 16>::zdc(c, W, dc, 1.f); break;
        default: fail(CTDM_EINVAL, "time_pass: unknown pass");
        }
    };
    launch(0);  // warm-up
    hipEvent_t e0, e1;
    CTDM_HIP(hipEventCreate(&e0));
    CTDM_HIP(hipEventCreate(&e1));
    CTDM_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; ++r) launch(r + 1);
    CTDM_HIP(hipEventRecord(e1, c.stream));
    CTDM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    CTDM_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return ms / reps;
}

// The positivity loop's own sequence (x, forward y, z + data consistency, inverse y) run `iters` times
// on the context's volume, each launch carrying its own start/end events (hipExtLaunchKernel): per-pass
// average durations in the context they run in (each pass reads what the previous one wrote), with the
// tuned choices.
// Stand-in data-consistency list for timing before any reconstruction at this V: `per` entries per
// pencil spread evenly over z (a 60-angle cell lands 20-30 samples on a typical pencil).
__global__ void k_dc_synth(DcList dc, int V, int per)
{
    const long long VV = (long long)V * V;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= VV) return;
    for (int k = 0; k < per; ++k) {
        dc.idx[p * dc.cap + k] = (short)((k * V) / per);
        dc.val[p * dc.cap + k] = make_float2(1e-3f, -1e-3f);
    }
    dc.cnt[p] = per;
}

template <int V>
static void time_loop_v(Ctx& c, int iters, double ms[3], double bytes[3])
{
    const size_t vol = (size_t)V * V * V, VV = (size_t)V * V;
    const std::string wname = "vol.W";
    const bool have = c.dc_V == V && c.scratch.count(wname) && c.scratch[wname].bytes >= sizeof(float2) * vol;
    const int per = std::min(24, V / 2);
    const int cap = have ? std::max(1, c.dc_cap) : per;
    float2* W = (float2*)c.buf(wname, sizeof(float2) * vol);
    const DcList dc = dc_alloc(c, V, cap);
    if (!have) {
        fill(c, W, 2 * (long long)vol, 1.f);
        hipLaunchKernelGGL(k_dc_synth, dim3(grid1d((long long)VV, 256)), dim3(256), 0, c.stream, dc, V, per);
        CTDM_CHECK_LAUNCH("k_dc_synth");
    }
    float2* nbuf = (float2*)c.buf("tp.n", sizeof(float2) * vol);
    const Proj p = make_proj(ewald_geom(ctdm_optics{532e-9, 1.337, 0.4e-6}, V));
    std::vector<int> h(VV);
    CTDM_HIP(hipMemcpy(h.data(), dc.cnt, sizeof(int) * VV, hipMemcpyDeviceToHost));
    double nent = 0;
    for (int v : h) nent += v;
    const double vb = 2.0 * sizeof(float2) * vol;
    bytes[0] = vb;
    bytes[1] = vb;
    bytes[2] = vb + sizeof(int) * (double)VV + nent * (sizeof(float2) + sizeof(short));
    const int nk = 4 * iters;
    std::vector<hipEvent_t> ev(2 * nk);
    for (auto& e : ev) CTDM_HIP(hipEventCreate(&e));
    Vol<V>::x(c, 1, W, nbuf, p, 1.f / (float)vol, true);  // warm-up of the sequence
    auto arm = [&](int k) {
        c.tev[0] = ev[2 * k];
        c.tev[1] = ev[2 * k + 1];
    };
    for (int it = 0; it < iters; ++it) {
        arm(4 * it);
        Vol<V>::x(c, 1, W, nbuf, p, 1.f / (float)vol, true);
        arm(4 * it + 1);
        Vol<V>::y(c, W, -1);
        arm(4 * it + 2);
        Vol<V>::zdc(c, W, dc);
        arm(4 * it + 3);
        Vol<V>::y(c, W, +1);
    }
    c.tev[0] = c.tev[1] = nullptr;
    CTDM_HIP(hipEventSynchronize(ev[2 * nk - 1]));
    double acc[3] = {0, 0, 0};
    for (int k = 0; k < nk; ++k) {
        float t = 0;
        CTDM_HIP(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
        const int kind = k % 4 == 0 ? 0 : (k % 4 == 2 ? 2 : 1);
        acc[kind] += t;
    }
    for (auto& e : ev) hipEventDestroy(e);
    ms[0] = acc[0] / iters;
    ms[1] = acc[1] / (2.0 * iters);
    ms[2] = acc[2] / iters;
}

void time_loop(Ctx& c, int V, int iters, double ms[3], double bytes[3])
{
    CTDM_REQUIRE(!c.capturing, "time_loop during capture");
    CTDM_REQUIRE(iters > 0 && ms && bytes, "time_loop: bad arguments");
    switch (V) {
    case 32: time_loop_v<32>(c, iters, ms, bytes); break;
    case 64: time_loop_v<64>(c, iters, ms, bytes); break;
    case 128: time_loop_v<128>(c, iters, ms, bytes); break;
    case 192: time_loop_v<192>(c, iters, ms, bytes); break;
    case 256: time_loop_v<256>(c, iters, ms, bytes); break;
    default: fail(CTDM_EUNSUPPORTED, "time_loop: volume side has no kernel");
    }
}

// Average ms of one gridding pass (gridding + hit averaging + list + inverse z-FFT) at side V with B
// pencils per workgroup, for A angles on an NA 1.0 cone of unit spectra (timing only; the pass's work per
// pencil does not depend on the spectrum's values).
template <int V>
static double time_grid_v(Ctx& c, int B, int A, int reps)
{
    const size_t VV = (size_t)V * V;
    const ctdm_optics o{532e-9, 1.337, 102.4e-6 / V};
    const EwaldGeom g = ewald_geom(o, V);
    std::vector<ctdm_angle> ang(A);
    const double s = 1.0 / 1.337;
    for (int a = 0; a < A; ++a) {
        const double az = 2.0 * M_PI * a / A;
        ang[a] = ctdm_angle{std::asin(s * std::sin(az)), std::asin(s * std::cos(az))};
    }
    const double* rows = upload_rows(c, "tg.rows", angle_rows(g, ang.data(), A));
    float2* spec = (float2*)c.buf("tg.spec", sizeof(float2) * VV * A);
    fill(c, spec, 2 * (long long)(VV * A), 1.f);
    float2* W = (float2*)c.buf("vol.W", sizeof(float2) * VV * V);
    const DcList dc = dc_alloc(c, V, A);
    auto launch = [&] {
        if (B == 8)
            VolOps<V, 8>::grid_pass(c, spec, rows, A, g, false, W, dc);
        else
            VolOps<V, 16>::grid_pass(c, spec, rows, A, g, false, W, dc);
    };
    launch();
    hipEvent_t e0, e1;
    CTDM_HIP(hipEventCreate(&e0));
    CTDM_HIP(hipEventCreate(&e1));
    CTDM_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; ++r) launch();
    CTDM_HIP(hipEventRecord(e1, c.stream));
    CTDM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    CTDM_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c.dc_V = 0;  // the list now holds this timing's entries, not a reconstruction's
    return ms / reps;
}

double time_grid(Ctx& c, int V, int B, int A, int reps)
{
    CTDM_REQUIRE(!c.capturing, "time_grid during capture");
    CTDM_REQUIRE((B == 8 || B == 16) && A > 0 && A < 32768 && reps > 0, "time_grid: bad arguments");
    switch (V) {
    case 32: return time_grid_v<32>(c, B, A, reps);
    case 64: return time_grid_v<64>(c, B, A, reps);
    case 128: return time_grid_v<128>(c, B, A, reps);
    case 192: return time_grid_v<192>(c, B, A, reps);
    case 256: return time_grid_v<256>(c, B, A, reps);
    default: fail(CTDM_EUNSUPPORTED, "time_grid: volume side has no kernel");
    }
}

// The y pass (8 pencils per workgroup) and the copy yardstick at V = 512: a 1 GiB volume, four times the
// 256 MB last-level cache, so these rates are HBM rates (at 256^3 the 128 MiB volume can stay cached
// between passes). Only for timing: no reconstruction runs at this size.
static double time_pass_512(Ctx& c, int kind, int reps, double* bytes)
{
    constexpr int V = 512;
    const size_t vol = (size_t)V * V * V;
    float2* W = (float2*)c.buf("tp.W512", sizeof(float2) * vol);
    float2* dst = kind == PASS_COPY ? (float2*)c.buf("tp.n512", sizeof(float2) * vol) : nullptr;
    fill(c, W, 2 * (long long)vol, 1.f);
    *bytes = 2.0 * sizeof(float2) * vol;
    const long long n4 = (long long)vol / 2;
    const int copy_grid = (int)std::min<long long>(grid1d(n4, 256), 8192);
    auto launch = [&](int r) {
        if (kind == PASS_COPY) {
            hipLaunchKernelGGL(k_copy_vol, dim3(copy_grid), dim3(256), 0, c.stream, (const float4*)W, (float4*)dst, n4);
            CTDM_CHECK_LAUNCH("k_copy_vol");
        } else {
            VolOps<V, 8>::y(c, W, (r & 1) ? 1 : -1);
        }
    };
    launch(0);
    hipEvent_t e0, e1;
    CTDM_HIP(hipEventCreate(&e0));
    CTDM_HIP(hipEventCreate(&e1));
    CTDM_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; ++r) launch(r + 1);
    CTDM_HIP(hipEventRecord(e1, c.stream));
    CTDM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    CTDM_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return ms / reps;
}

double time_pass(Ctx& c, int kind, int V, int B, int reps, double* bytes)
{
    CTDM_REQUIRE(!c.capturing, "time_pass during capture");
    CTDM_REQUIRE(reps > 0 && (B == 8 || B == 16), "time_pass: bad arguments");
    CTDM_REQUIRE(kind >= 0 && kind <= PASS_COPY && kind != PASS_GRID,
                 "time_pass: pass must be 0 (x), 1 (y), 2 (z+dc) or 4 (plain copy)");
    if (V == 512) {
        CTDM_REQUIRE(kind == PASS_Y || kind == PASS_COPY, "time_pass: at V = 512 only the y pass and the copy");
        return time_pass_512(c, kind, reps, bytes);
    }
    switch (V) {
    case 32: return time_pass_v<32>(c, kind, B, reps, bytes);
    case 64: return time_pass_v<64>(c, kind, B, reps, bytes);
    case 128: return time_pass_v<128>(c, kind, B, reps, bytes);
    case 192: return time_pass_v<192>(c, kind, B, reps, bytes);
    case 256: return time_pass_v<256>(c, kind, B, reps, bytes);
    default: fail(CTDM_EUNSUPPORTED, "time_pass: volume side has no kernel");
    }
}

}  // namespace ctdm
