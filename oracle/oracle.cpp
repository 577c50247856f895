// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// FP64 host restatement of the reference cTDM reconstruction path (y-h-Lin/cTDM), used as the
// correctness checker of libctdm. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load liboracle.so; the product path never calls it.
//
// Each function cites the reference code it restates. Layouts follow the reference:
// idx = x + W*y (+ W*H*z), complex as interleaved (re, im) doubles.
//
// What pins each part (checked in tests/test_oracle_*.py):
//   orc_fft                 numpy.fft on random data; brute-force DFT for n <= 64
//   orc_dct_chain stages    the reference's own dumps src/Laplace/DCT/DCT2/IDCT.256.256.raw
//                           (DCT_Unwrapping.cu:35,47,51,55); scipy dctn/idctn (ortho) for the exact
//                           transforms; LS closed form unwrap(wrap(phi)) = phi - mean(phi) for
//                           smooth phi
//   orc_retrieve            closed forms: pure carrier -> phase 0, log-amplitude 0; a known smooth
//                           phase/amplitude is recovered; brute-force cropped DFT at small H
//   orc_find_sideband       carrier position (0.25, 0.25) -> (768, 768), radius 85
//   orc_goldstein           hand-enumerated residue / cut / fill cases; integer invariants
//   orc_ewald_* / recon     single-plane-wave closed form; scatter/gather adjointness; the
//                           est_n / modify_F identity on feasible n; idempotent positivity
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;
static const double PI = 3.14159265358979323846;

// ---------------------------------------------------------------- FFT (FP64, any length)
// Recursive mixed-radix decimation in time; the reference uses cuFFT C2C (FFT.cu:636-838):
// forward exp(-2 pi i jk/n) unnormalised, inverse exp(+2 pi i jk/n) unnormalised.
namespace {

int smallest_factor(int n)
{
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

struct Twiddles {
    std::vector<cd> w;  // exp(-2 pi i k / n)
};

const std::vector<cd>& twiddles(int n)
{
    static std::mutex mu;
    static std::map<int, Twiddles> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto& t = cache[n];
    if ((int)t.w.size() != n) {
        t.w.resize(n);
        for (int k = 0; k < n; ++k) t.w[k] = std::polar(1.0, -2.0 * PI * k / n);
    }
    return t.w;
}

// out[k] = sum_j in[j*stride] w^(jk), w = exp(dir 2 pi i / n); `tw` is the table of the top-level n0
// and `tstep` = n0 / n.
void fft_rec(const cd* in, cd* out, int n, long stride, int dir, const std::vector<cd>& tw, int tstep)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int p = smallest_factor(n), m = n / p;
    std::vector<cd> sub(n);
    for (int r = 0; r < p; ++r) fft_rec(in + r * stride, sub.data() + r * m, m, stride * p, dir, tw, tstep * p);
    const int n0 = (int)tw.size();
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            const int kk = k + q * m;
            cd acc = 0;
            for (int r = 0; r < p; ++r) {
                const long e = ((long)r * kk % n) * tstep % n0;
                const cd w = dir < 0 ? tw[e] : std::conj(tw[e]);
                acc += w * sub[r * m + k];
            }
            out[kk] = acc;
        }
    }
}

void fft1(cd* data, int n, long stride, int dir)
{
    const auto& tw = twiddles(n);
    std::vector<cd> in(n), out(n);
    for (int i = 0; i < n; ++i) in[i] = data[i * stride];
    fft_rec(in.data(), out.data(), n, 1, dir, tw, 1);
    for (int i = 0; i < n; ++i) data[i * stride] = out[i];
}

// 2-D transform of a W x H image (rows then columns).
void fft2(cd* d, int W, int H, int dir)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) fft1(d + (long)y * W, W, 1, dir);
#pragma omp parallel for schedule(static)
    for (int x = 0; x < W; ++x) fft1(d + x, H, W, dir);
}

// 3-D transform of a V^3 volume.
void fft3(cd* d, int V, int dir)
{
    const long VV = (long)V * V;
#pragma omp parallel for schedule(static)
    for (long r = 0; r < VV; ++r) fft1(d + r * V, V, 1, dir);
#pragma omp parallel for schedule(static)
    for (long r = 0; r < VV; ++r) {
        const long x = r % V, z = r / V;
        fft1(d + x + z * VV, V, V, dir);
    }
#pragma omp parallel for schedule(static)
    for (long r = 0; r < VV; ++r) fft1(d + r, V, VV, dir);
}

// cuFFT2Dshift (FFT.cu:220-236): quadrant swap, even sizes.
void shift2(cd* d, int W, int H)
{
    for (int j = 0; j < H / 2; ++j)
        for (int i = 0; i < W / 2; ++i) {
            std::swap(d[i + (long)W * j], d[(i + W / 2) + (long)W * (j + H / 2)]);
            std::swap(d[(i + W / 2) + (long)W * j], d[i + (long)W * (j + H / 2)]);
        }
}

// cuFFT3Dshift (FFT.cu:324-350): octant swap, even V.
void shift3(cd* d, int V)
{
    const int h = V / 2;
    const long VV = (long)V * V;
    auto at = [&](int x, int y, int z) -> cd& { return d[x + (long)V * y + VV * z]; };
    for (int k = 0; k < h; ++k)
        for (int j = 0; j < h; ++j)
            for (int i = 0; i < h; ++i) {
                std::swap(at(i, j, k), at(i + h, j + h, k + h));
                std::swap(at(i, j + h, k), at(i + h, j, k + h));
                std::swap(at(i + h, j, k), at(i, j + h, k + h));
                std::swap(at(i + h, j + h, k), at(i, j, k + h));
            }
}

double wrap_pi(double a)  // to (-pi, pi]
{
    double w = a - 2.0 * PI * std::floor(a / (2.0 * PI) + 0.5);
    if (w <= -PI) w += 2.0 * PI;
    if (w > PI) w -= 2.0 * PI;
    return w;
}

}  // namespace

extern "C" {

int orc_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// In-place 1-D transform of `count` contiguous interleaved complex rows of length n.
void orc_fft(double* data, int n, int count, int dir)
{
    cd* d = reinterpret_cast<cd*>(data);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < count; ++r) fft1(d + (long)r * n, n, 1, dir);
}

// In-place 3-D transform of a V^3 volume; dir > 0 additionally scales by 1/V^3 (cuFFT3D(+1) then
// scaleFFT3D, FFT.cu:807-838, 900-916: no conjugation).
void orc_fft3d(double* data, int V, int dir)
{
    cd* d = reinterpret_cast<cd*>(data);
    fft3(d, V, dir);
    if (dir > 0) {
        const double s = 1.0 / ((double)V * V * V);
        const long n = (long)V * V * V;
        for (long i = 0; i < n; ++i) d[i] *= s;
    }
}

// ---------------------------------------------------------------- retrieval (Algorithm A)
// obtainRadius (HilbertTransform.cu:3425-3455) on the fftshifted background spectrum
// (HilbertTransform.cu:1564-1572): argmax of log10|F| over rows [row_start, H), raster order,
// strict '>' (first maximum wins); radius = |r - H/2| / 3 (integer division).
void orc_find_sideband(const float* bg, int H, int W, int row_start, int* row, int* col, int* radius)
{
    std::vector<cd> B((size_t)H * W);
    for (size_t i = 0; i < B.size(); ++i) B[i] = cd(bg[i], 0.0);
    fft2(B.data(), W, H, -1);
    shift2(B.data(), W, H);
    if (row_start <= 0) row_start = 600 * H / 1024;
    double best = -1e300;
    int r = 0, c = 0;
    for (int j = row_start; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            const double v = std::log10(std::abs(B[i + (size_t)W * j]));
            if (v > best) {
                best = v;
                r = j;
                c = i;
            }
        }
    *row = r;
    *col = c;
    *radius = std::abs(r - H / 2) / 3;
}

// extractQPI_AlgorithmA (HilbertTransform.cu:1546-1585) + calcWrapPhase (:3618-3639):
// FFT2 -> cuFFT2Dshift -> get1stOrder_new (:3566-3586, M-generalised) -> cuFFT2Dshift (M) ->
// inverse FFT2 -> scaleFFT2D (x 1/M^2 and conjugate, FFT.cu:853-868) ->
//   phase_ref = atan2(BG) - atan2(SP) in (-2 pi, 2 pi), amp = |SP| / |BG| (non-finite -> 0).
// Also returns wrapped = phase_ref mod 2 pi in (-pi, pi], logamp = ln(amp), and the unconjugated
// ratio field SP/BG (interleaved). Any output pointer may be NULL.
void orc_retrieve(const float* holo, const float* bg, int H, int W, int row, int col, int radius,
                  int M, double* phase_ref, double* amp, double* wrapped, double* logamp,
                  double* field)
{
    std::vector<cd> S((size_t)H * W), B((size_t)H * W);
    for (size_t i = 0; i < S.size(); ++i) {
        S[i] = cd(holo[i], 0.0);
        B[i] = cd(bg[i], 0.0);
    }
    fft2(S.data(), W, H, -1);
    fft2(B.data(), W, H, -1);
    shift2(S.data(), W, H);
    shift2(B.data(), W, H);
    std::vector<cd> S2((size_t)M * M), B2((size_t)M * M);
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            const int ii = i + (col - M / 2), jj = j + (row - M / 2);
            const bool in = ii >= col - radius && ii <= col + radius && jj >= row - radius && jj <= row + radius;
            S2[i + (size_t)M * j] = in ? S[ii + (size_t)W * jj] : cd(0, 0);
            B2[i + (size_t)M * j] = in ? B[ii + (size_t)W * jj] : cd(0, 0);
        }
    shift2(S2.data(), M, M);
    shift2(B2.data(), M, M);
    fft2(S2.data(), M, M, +1);
    fft2(B2.data(), M, M, +1);
    const double sc = 1.0 / ((double)M * M);
    for (size_t i = 0; i < S2.size(); ++i) {
        S2[i] = std::conj(S2[i] * sc);
        B2[i] = std::conj(B2[i] * sc);
    }
    for (size_t i = 0; i < S2.size(); ++i) {
        double p = std::atan2(B2[i].imag(), B2[i].real()) - std::atan2(S2[i].imag(), S2[i].real());
        if (!std::isfinite(p)) p = 0;
        double a = std::abs(S2[i]) / std::abs(B2[i]);
        if (!std::isfinite(a)) a = 0;
        if (phase_ref) phase_ref[i] = p;
        if (amp) amp[i] = a;
        if (wrapped) wrapped[i] = wrap_pi(p);
        if (logamp) {
            double l = std::log(a);
            logamp[i] = std::isfinite(l) ? l : 0.0;
        }
        if (field) {
            const cd f = std::conj(S2[i]) / std::conj(B2[i]);
            field[2 * i] = std::isfinite(f.real()) ? f.real() : 0.0;
            field[2 * i + 1] = std::isfinite(f.imag()) ? f.imag() : 0.0;
        }
    }
}

// ---------------------------------------------------------------- DCT least-squares unwrapping
// DCT_UWLS_Unwrapped (DCT_Unwrapping.cu:10-59) and its kernels.
enum : unsigned { SINCOS = 1u, REF_EIG = 2u, REF_IDCT = 4u, CONGRUENT = 8u };

static int sgn(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }

// Wrap used by NumericDerivative1 (DCT_Unwrapping.cu:378-379):
//   t - 2 pi sign(t) floor(|pi sign(t) + t| / 2 pi)
static double wrapdiff(double t)
{
    const int s = sgn(t);
    return t - 2.0 * PI * s * std::floor(std::fabs(PI * s + t) / (2.0 * PI));
}

// RHS of the Poisson problem.
//   WRAPDIFF: NumericDerivative1 + SumDerivative (DCT_Unwrapping.cu:336-358, 394-417)
//   SINCOS:   cos(phi) L[sin phi] - sin(phi) L[cos phi], L the same zero-flux 5-point Laplacian
//             (the trigonometric RHS of FastUnwrapping.cu:37-96, discretised consistently with
//             the DCT-II solver).
static void laplacian_rhs(const double* src, int N, unsigned flags, double* out)
{
    std::vector<double> dx((size_t)N * N), dy((size_t)N * N);
    if (flags & SINCOS) {
        std::vector<double> s((size_t)N * N), c((size_t)N * N);
        for (size_t i = 0; i < s.size(); ++i) {
            s[i] = std::sin(src[i]);
            c[i] = std::cos(src[i]);
        }
        auto L = [&](const std::vector<double>& g, int i, int j) {
            const size_t k = i + (size_t)N * j;
            double v = 0;
            if (i < N - 1) v += g[k + 1] - g[k];
            if (i > 0) v -= g[k] - g[k - 1];
            if (j < N - 1) v += g[k + N] - g[k];
            if (j > 0) v -= g[k] - g[k - N];
            return v;
        };
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                const size_t k = i + (size_t)N * j;
                out[k] = c[k] * L(s, i, j) - s[k] * L(c, i, j);
            }
        return;
    }
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const size_t k = i + (size_t)N * j;
            const double tx = (i == N - 1) ? -src[k] : src[k + 1] - src[k];
            const double ty = (j == N - 1) ? -src[k] : src[k + N] - src[k];
            dx[k] = wrapdiff(tx);
            dy[k] = wrapdiff(ty);
        }
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const size_t k = i + (size_t)N * j;
            const double c1 = (i == N - 1) ? 0 : dx[k];
            const double c2 = (i == 0) ? 0 : dx[k - 1];
            const double c3 = (j == N - 1) ? 0 : dy[k];
            const double c4 = (j == 0) ? 0 : dy[k - N];
            out[k] = (c1 - c2) + (c3 - c4);
        }
}

static void transpose(const double* in, double* out, int N)
{
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) out[j + (size_t)N * i] = in[i + (size_t)N * j];
}

// One pass of myDCT (DCT_Unwrapping.cu:61-83): per row, float2cufft mirror to 2N (:144-162),
// forward 2N-point FFT, countFDCT (:202-216) with dTemp = 1/sqrt(2N).
static void dct_rows(const double* in, double* out, int N)
{
    const double dT = 1.0 / std::sqrt(2.0 * N);
    std::vector<cd> y(2 * N);
    for (int j = 0; j < N; ++j) {
        for (int i = 0; i < N; ++i) {
            y[i] = cd(in[i + (size_t)N * j], 0);
            y[2 * N - 1 - i] = y[i];
        }
        fft1(y.data(), 2 * N, 1, -1);
        for (int i = 0; i < N; ++i) {
            const double C = y[i].real() * std::cos(i * PI / (2.0 * N));
            const double S = y[i].imag() * std::sin(i * PI / (2.0 * N));
            out[i + (size_t)N * j] = (i == 0) ? y[i].real() * dT / std::sqrt(2.0) : (C + S) * dT;
        }
    }
}

// One pass of myIDCT (DCT_Unwrapping.cu:85-107): per row, countIDCT (:218-246) builds the
// Hermitian 2N completion, inverse 2N-point FFT, cufft2float (:164-177) keeps Re of the first N.
// The reference's right half uses angle (i+1)*pi/2N (REF_IDCT); the exact inverse uses i*pi/2N.
static void idct_rows(const double* in, double* out, int N, bool refcompat)
{
    const double dT = 1.0 / std::sqrt(2.0 * N);
    const double d = refcompat ? 1.0 : 0.0;
    std::vector<cd> z(2 * N);
    for (int j = 0; j < N; ++j) {
        const double* row = in + (size_t)N * j;
        for (int i = 0; i < N; ++i) {
            const double C = std::cos(i * PI / (2.0 * N)), S = std::sin(i * PI / (2.0 * N));
            const double V = row[i] * dT;
            z[i] = (i == 0) ? cd(C * V * std::sqrt(2.0), S * V * std::sqrt(2.0)) : cd(C * V, S * V);
            if (i == 0) {
                z[N] = 0;
            } else {
                const double S2 = std::sin((i + d) * PI / (2.0 * N));
                const double C2 = -std::cos((i + d) * PI / (2.0 * N));
                const double V2 = row[N - i] * dT;
                z[N + i] = cd(S2 * V2, C2 * V2);
            }
        }
        fft1(z.data(), 2 * N, 1, +1);
        for (int i = 0; i < N; ++i) out[i + (size_t)N * j] = z[i].real();
    }
}

// devConstant (DCT_Unwrapping.cu:300-312): divide by the Laplacian eigenvalue, DC := 0.
//   exact:   lambda_ij = -4 (sin^2(pi i / 2N) + sin^2(pi j / 2N))
//   REF_EIG: lambda_ij = 2 cos(4 pi i / (4N - 1)) + 2 cos(4 pi j / (4N - 1)) - 4
static void poisson_divide(double* buf, int N, bool refcompat)
{
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const size_t k = i + (size_t)N * j;
            if (i == 0 && j == 0) {
                buf[k] = 0;
                continue;
            }
            double lam;
            if (refcompat)
                lam = 2 * std::cos(4 * PI * i / (4.0 * N - 1)) + 2 * std::cos(4 * PI * j / (4.0 * N - 1)) - 4;
            else {
                const double a = std::sin(PI * i / (2.0 * N)), b = std::sin(PI * j / (2.0 * N));
                lam = -4.0 * (a * a + b * b);
            }
            buf[k] /= lam;
        }
}

// Stages of the DCT unwrapper on an N x N map. `stage_in` selects where the chain starts
// (0: wrapped phase, 1: Laplacian, 2: DCT coefficients, 3: divided coefficients) so each stage can
// be checked against the reference dump of its own input. Outputs may be NULL.
void orc_dct_chain(const double* in, int N, unsigned flags, int stage_in, double* lap, double* dct,
                   double* dct2, double* out)
{
    const size_t n = (size_t)N * N;
    std::vector<double> L(n), D(n), D2(n), t(n), t2(n), U(n);
    if (stage_in <= 0)
        laplacian_rhs(in, N, flags, L.data());
    else if (stage_in == 1)
        std::memcpy(L.data(), in, n * sizeof(double));
    if (stage_in <= 1) {
        dct_rows(L.data(), t.data(), N);
        transpose(t.data(), t2.data(), N);
        dct_rows(t2.data(), t.data(), N);
        transpose(t.data(), D.data(), N);
    } else if (stage_in == 2) {
        std::memcpy(D.data(), in, n * sizeof(double));
    }
    if (stage_in <= 2) {
        D2 = D;
        poisson_divide(D2.data(), N, flags & REF_EIG);
    } else {
        std::memcpy(D2.data(), in, n * sizeof(double));
    }
    transpose(D2.data(), t.data(), N);
    idct_rows(t.data(), t2.data(), N, flags & REF_IDCT);
    transpose(t2.data(), t.data(), N);
    idct_rows(t.data(), U.data(), N, flags & REF_IDCT);
    if ((flags & CONGRUENT) && stage_in <= 0) {
        for (size_t k = 0; k < n; ++k) U[k] = in[k] + 2.0 * PI * std::round((U[k] - in[k]) / (2.0 * PI));
    }
    if (lap) std::memcpy(lap, L.data(), n * sizeof(double));
    if (dct) std::memcpy(dct, D.data(), n * sizeof(double));
    if (dct2) std::memcpy(dct2, D2.data(), n * sizeof(double));
    if (out) std::memcpy(out, U.data(), n * sizeof(double));
}

// Batched convenience: unwrap `batch` maps.
void orc_unwrap_dct(const double* in, double* out, int batch, int N, unsigned flags)
{
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < batch; ++b)
        orc_dct_chain(in + (size_t)b * N * N, N, flags, 0, nullptr, nullptr, nullptr,
                      out + (size_t)b * N * N);
}

}  // extern "C"
