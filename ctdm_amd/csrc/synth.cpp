// Seeded synthetic inputs for the cTDM path (host C++). The reference reads BMP holograms from
// disk (HilbertTransform.cu:2680-2731); these generators produce the same kind of data in memory.
//
// Random numbers are counter-based (SplitMix64 of seed, stream and index), so every pixel is
// reproducible independently of threading.
//
// Models
//   hologram   I = A^2 + 1 + 2 A cos(2 pi (fx x + fy y) + phi) + N(0, sigma^2); background A = 1,
//              phi = 0 with independent noise.
//   cell       an ellipsoid (n_cell + i kappa) centred at the origin holding spherical nucleoli
//              (n_cell + (n_nucleoli - n_cell)) in its inner half, seen under A illumination directions
//              d_a on a cone n_med sin(theta) = NA_ill, azimuth 2 pi a / A, in image coordinates
//              (x along a row, y across rows, z along the optical axis). The reference transposes the
//              maps before gridding (cuCopyArray, Recon_kernel.cu:162) and takes f_x0 from AngY
//              (:175), so AngY is the tilt along image y: ang_y = asin(d_y), ang_x = asin(d_x).
//              Two field models (ctdm_synth_params.model):
//     1 (default) first-order Rytov. With o(r) = k0^2 (n(r)^2 - n_med^2), the Rytov field in the plane
//              z = 0 has the 2-D spectrum Psi(q) = i O(K) / (4 pi f_z), K = (q, f_z - f_z0),
//              f_z = sqrt(f_m^2 - |q + f_in|^2) (zero for evanescent q), where O is the analytic 3-D
//              Fourier transform of the phantom's potential (each body a uniform ellipsoid:
//              V 3 (sin u - u cos u) / u^3, u = 2 pi |S K|, times exp(-2 pi i K.c)). This is the
//              first-order Born scattering integral with the Green's function e^{ikR} / (4 pi R)
//              expanded in plane waves (Weyl); psi(r) = sum_q Psi(q) e^{2 pi i q.r} df^2 on the H x H
//              grid with the origin at the centre pixel; the sample field is exp(psi).
//     0        straight-ray projections (the geometric-optics limit of the Rytov phase):
//                 phi_a(x, y) = k0 * integral (Re n - n_med) dt,  ln A_a = -k0 * integral Im n dt
//              along d_a through (x, y, 0).
//   phantom    the same refractive index sampled at the voxel centres of the reconstruction grid:
//              volume x is image y, volume y image x, voxel (V/2, V/2, V/2) the origin (the
//              octant-shifted output of the pipeline).
//   wrapped    sum of 8 Gaussians (|peak| <= 40 rad, sigma 20-80 px) + U(-0.8 pi, 0.8 pi), wrapped
//              to (-pi, pi].
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/ctdm.h"

namespace {

inline uint64_t splitmix(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

inline uint64_t key(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return splitmix(splitmix(seed * 0x100000001b3ull ^ stream) + idx);
}

// Uniform in (0, 1].
inline double uni(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return ((key(seed, stream, idx) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

inline double gauss(uint64_t seed, uint64_t stream, uint64_t idx)
{
    const double u1 = uni(seed, stream, 2 * idx), u2 = uni(seed, stream, 2 * idx + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

template <class F>
void parallel_rows(int n, F f)
{
    const int nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) f(i);
        });
    for (auto& x : th) x.join();
}

struct Ellipsoid {
    double c[3], s[3];
};

// Chord length of the line p0 + t d (|d| = 1) through an axis-aligned ellipsoid.
inline double chord(const Ellipsoid& e, const double p0[3], const double d[3])
{
    double uu = 0, uw = 0, ww = 0;
    for (int i = 0; i < 3; ++i) {
        const double u = (p0[i] - e.c[i]) / e.s[i], w = d[i] / e.s[i];
        uu += u * u;
        uw += u * w;
        ww += w * w;
    }
    const double disc = uw * uw - ww * (uu - 1.0);
    return disc > 0 ? 2.0 * std::sqrt(disc) / ww : 0.0;
}

const double kPi = 3.14159265358979323846;

inline bool inside(const Ellipsoid& e, const double r[3])
{
    double q = 0;
    for (int i = 0; i < 3; ++i) {
        const double u = (r[i] - e.c[i]) / e.s[i];
        q += u * u;
    }
    return q <= 1.0;
}

// The cell phantom in micrometres and its potential contrasts o = k0^2 (n^2 - n_med^2): the cell body
// (n_cell + i kappa) and, inside it, each nucleolus adding k0^2 ((n_c + dn)^2 - n_c^2), n_c = n_cell + i kappa.
struct Phantom {
    Ellipsoid cell;
    std::vector<Ellipsoid> nuc;
    std::complex<double> o_cell, o_nuc;
};

Phantom make_phantom(uint64_t seed, const ctdm_synth_params& p)
{
    const double um = 1e6;
    Phantom ph;
    ph.cell = Ellipsoid{{0, 0, 0}, {p.semi_axes[0] * um, p.semi_axes[1] * um, p.semi_axes[2] * um}};
    ph.nuc.resize(p.n_nuc);
    for (int i = 0; i < p.n_nuc; ++i) {
        // uniform in the inner half of the cell (keeps nucleoli inside the ellipsoid)
        double v[3];
        do {
            for (int a = 0; a < 3; ++a) v[a] = 2.0 * uni(seed, 100 + i, a) - 1.0;
            seed += 0x51ed27;
        } while (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > 1.0);
        for (int a = 0; a < 3; ++a) {
            ph.nuc[i].c[a] = 0.5 * v[a] * ph.cell.s[a];
            ph.nuc[i].s[a] = p.r_nuc * um;
        }
    }
    const double k0 = 2.0 * kPi / (p.wavelength * um);
    const std::complex<double> nc(p.n_cell, p.kappa_cell);
    const std::complex<double> nn = nc + (p.n_nucleoli - p.n_cell);
    ph.o_cell = k0 * k0 * (nc * nc - p.n_med * p.n_med);
    ph.o_nuc = k0 * k0 * (nn * nn - nc * nc);
    return ph;
}

// 3-D Fourier transform of the indicator of a uniform ellipsoid at K [cycles / um].
std::complex<double> ellipsoid_ft(const Ellipsoid& e, const double K[3])
{
    double q = 0, kc = 0;
    for (int i = 0; i < 3; ++i) {
        q += (e.s[i] * K[i]) * (e.s[i] * K[i]);
        kc += K[i] * e.c[i];
    }
    const double u = 2.0 * kPi * std::sqrt(q);
    const double vol = 4.0 / 3.0 * kPi * e.s[0] * e.s[1] * e.s[2];
    const double shape = u < 1e-3 ? 1.0 - u * u / 10.0 : 3.0 * (std::sin(u) - u * std::cos(u)) / (u * u * u);
    return vol * shape * std::exp(std::complex<double>(0.0, -2.0 * kPi * kc));
}

std::complex<double> potential_ft(const Phantom& ph, const double K[3])
{
    std::complex<double> o = ph.o_cell * ellipsoid_ft(ph.cell, K);
    for (const auto& e : ph.nuc) o += ph.o_nuc * ellipsoid_ft(e, K);
    return o;
}

// In-place unnormalised inverse transform (exp(+2 pi i jk / n)) of n contiguous points, n a power of two.
void fft_line(std::complex<double>* x, int n, const std::vector<std::complex<double>>& w)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int step = n / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                const std::complex<double> t = w[(size_t)k * step] * x[i + k + len / 2];
                x[i + k + len / 2] = x[i + k] - t;
                x[i + k] += t;
            }
    }
}

// Unnormalised inverse 2-D transform of an n x n array (rows, then columns through a transpose).
void fft2_inverse(std::complex<double>* a, int n)
{
    std::vector<std::complex<double>> w(n / 2);
    for (int k = 0; k < n / 2; ++k) w[k] = std::polar(1.0, 2.0 * kPi * k / n);
    parallel_rows(n, [&](int r) { fft_line(a + (size_t)r * n, n, w); });
    std::vector<std::complex<double>> t((size_t)n * n);
    parallel_rows(n, [&](int r) {
        for (int c = 0; c < n; ++c) t[(size_t)c * n + r] = a[(size_t)r * n + c];
    });
    parallel_rows(n, [&](int r) { fft_line(t.data() + (size_t)r * n, n, w); });
    parallel_rows(n, [&](int r) {
        for (int c = 0; c < n; ++c) a[(size_t)c * n + r] = t[(size_t)r * n + c];
    });
}

}  // namespace

extern "C" {

void ctdm_synth_default_params(ctdm_synth_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->H = 1024;
    p->pixel = 0.1e-6;
    p->wavelength = 532e-9;
    p->n_med = 1.337;
    p->na_ill = 1.0;
    p->n_cell = 1.36;
    p->n_nucleoli = 1.39;
    p->kappa_cell = 2e-4;
    p->semi_axes[0] = 6e-6;
    p->semi_axes[1] = 5e-6;
    p->semi_axes[2] = 4e-6;
    p->n_nuc = 5;
    p->r_nuc = 1.0e-6;
    p->noise = 0.01;
    p->carrier[0] = 0.25;
    p->carrier[1] = 0.25;
    p->model = 1;
}

ctdm_status ctdm_synth_hologram(uint64_t seed, int H, float* holo, float* bg)
{
    if (!holo || !bg || H < 16) return CTDM_EINVAL;
    const double c = H / 2.0, r = 80.0 * H / 1024.0;
    parallel_rows(H, [&](int y) {
        for (int x = 0; x < H; ++x) {
            const double rr = ((x - c) * (x - c) + (y - c) * (y - c)) / (r * r);
            const double u = (x - c) / c, v = (y - c) / c;
            const double phi = 12.0 * std::sqrt(std::max(0.0, 1.0 - rr)) +
                               (1.5 * u + 1.0 * v + 0.8 * u * u - 0.6 * u * v + 0.5 * v * v);
            const double carrier = 2.0 * kPi * (0.25 * x + 0.25 * y);
            const uint64_t idx = (uint64_t)y * H + x;
            holo[idx] = (float)(2.0 + 2.0 * std::cos(carrier + phi) + 0.01 * gauss(seed, 1, idx));
            bg[idx] = (float)(2.0 + 2.0 * std::cos(carrier) + 0.01 * gauss(seed, 2, idx));
        }
    });
    return CTDM_OK;
}

ctdm_status ctdm_synth_cell(uint64_t seed, int A, const ctdm_synth_params* p, float* holo, float* bg,
                            ctdm_angle* angles)
{
    if (!p || !holo || !bg || !angles || A <= 0 || p->H < 16 || (p->H & (p->H - 1))) return CTDM_EINVAL;
    if (p->model != 0 && p->model != 1) return CTDM_EINVAL;
    const int H = p->H;
    const Phantom ph = make_phantom(seed, *p);
    const double um = 1e6;  // work in micrometres
    const double pix = p->pixel * um, lam = p->wavelength * um, k0 = 2.0 * kPi / lam;
    const double sin_t = p->na_ill / p->n_med;
    const double cos_t = std::sqrt(std::max(0.0, 1.0 - sin_t * sin_t));
    std::vector<std::array<double, 3>> dir(A);
    for (int a = 0; a < A; ++a) {
        const double az = 2.0 * kPi * a / A;
        dir[a] = {sin_t * std::cos(az), sin_t * std::sin(az), cos_t};
        angles[a].ang_x = std::asin(dir[a][0]);
        angles[a].ang_y = std::asin(dir[a][1]);
    }
    // log-amplitude and phase of every angle, [a][y][x]
    std::vector<double> la((size_t)A * H * H), phi((size_t)A * H * H);
    if (p->model == 0) {
        const double dn_cell = p->n_cell - p->n_med, dn_nuc = p->n_nucleoli - p->n_cell;
        parallel_rows(A * H, [&](int ay) {
            const int a = ay / H, y = ay % H;
            const double* d = dir[a].data();
            for (int x = 0; x < H; ++x) {
                const double p0[3] = {(x - H / 2) * pix, (y - H / 2) * pix, 0.0};
                double f = 0.0, l = 0.0;
                const double lc = chord(ph.cell, p0, d);
                if (lc > 0) {
                    f = k0 * dn_cell * lc;
                    l = -k0 * p->kappa_cell * lc;
                    for (const auto& e : ph.nuc) f += k0 * dn_nuc * chord(e, p0, d);
                }
                const size_t i = ((size_t)a * H + y) * H + x;
                phi[i] = f;
                la[i] = l;
            }
        });
    } else {
        const double fm = p->n_med / lam, dfq = 1.0 / (H * pix);
        std::vector<std::complex<double>> psi((size_t)H * H);
        for (int a = 0; a < A; ++a) {
            const double fin[3] = {fm * dir[a][0], fm * dir[a][1], fm * dir[a][2]};
            // Psi(q) (-1)^(u+v) on the DFT grid (the sign factor puts the origin at the centre pixel)
            parallel_rows(H, [&](int v) {
                const int sv = v < H / 2 ? v : v - H;
                for (int u = 0; u < H; ++u) {
                    const int su = u < H / 2 ? u : u - H;
                    const double qx = su * dfq, qy = sv * dfq;
                    const double fx = qx + fin[0], fy = qy + fin[1];
                    const double s2 = fm * fm - fx * fx - fy * fy;
                    std::complex<double> val = 0.0;
                    if (s2 > 0) {
                        const double fz = std::sqrt(s2);
                        const double K[3] = {qx, qy, fz - fin[2]};
                        const std::complex<double> O = potential_ft(ph, K);
                        val = std::complex<double>(0.0, 1.0) * O / (4.0 * kPi * fz) * (dfq * dfq);
                        if ((u + v) & 1) val = -val;
                    }
                    psi[(size_t)v * H + u] = val;
                }
            });
            fft2_inverse(psi.data(), H);
            for (size_t i = 0; i < (size_t)H * H; ++i) {
                la[(size_t)a * H * H + i] = psi[i].real();
                phi[(size_t)a * H * H + i] = psi[i].imag();
            }
        }
    }
    parallel_rows(A * H, [&](int ay) {
        const int a = ay / H, y = ay % H;
        float* hp = holo + (size_t)a * H * H + (size_t)y * H;
        float* bp = bg + (size_t)a * H * H + (size_t)y * H;
        for (int x = 0; x < H; ++x) {
            const uint64_t idx = ((uint64_t)a * H + y) * H + x;
            const double Aamp = std::exp(la[idx]);
            const double carrier = 2.0 * kPi * (p->carrier[0] * x + p->carrier[1] * y);
            // sample field E = A exp(i phi) (+ complex noise); I = |E|^2 + 1 + 2 Re(E exp(i carrier))
            double er = Aamp * std::cos(phi[idx]), ei = Aamp * std::sin(phi[idx]);
            if (p->field_noise > 0) {
                const double s = p->field_noise / std::sqrt(2.0);
                er += s * gauss(seed, 6, idx);
                ei += s * gauss(seed, 7, idx);
            }
            const double cc = std::cos(carrier), ss = std::sin(carrier);
            hp[x] = (float)(er * er + ei * ei + 1.0 + 2.0 * (er * cc - ei * ss) + p->noise * gauss(seed, 3, idx));
            bp[x] = (float)(2.0 + 2.0 * std::cos(carrier) + p->noise * gauss(seed, 4, idx));
        }
    });
    return CTDM_OK;
}

ctdm_status ctdm_synth_phantom(uint64_t seed, const ctdm_synth_params* p, int V, double dx, ctdm_complex* n)
{
    if (!p || !n || V <= 0 || !(dx > 0)) return CTDM_EINVAL;
    const Phantom ph = make_phantom(seed, *p);
    const double h = dx * 1e6;
    parallel_rows(V, [&](int z) {
        for (int yv = 0; yv < V; ++yv)
            for (int xv = 0; xv < V; ++xv) {
                const double r[3] = {(yv - V / 2) * h, (xv - V / 2) * h, (z - V / 2) * h};  // image x, y, z
                double re = p->n_med, im = 0.0;
                if (inside(ph.cell, r)) {
                    re = p->n_cell;
                    im = p->kappa_cell;
                    for (const auto& e : ph.nuc)
                        if (inside(e, r)) re += p->n_nucleoli - p->n_cell;
                }
                ctdm_complex& o = n[((size_t)z * V + yv) * V + xv];
                o.x = (float)re;
                o.y = (float)im;
            }
    });
    return CTDM_OK;
}

ctdm_status ctdm_synth_wrapped(uint64_t seed, int batch, int H, int W, float* out)
{
    if (!out || batch < 0 || H < 2 || W < 2) return CTDM_EINVAL;
    parallel_rows(batch * H, [&](int by) {
        const int b = by / H, y = by % H;
        double g[8][4];
        for (int i = 0; i < 8; ++i) {
            g[i][0] = (2.0 * uni(seed, 200 + b, 4 * i) - 1.0) * 40.0;  // peak
            g[i][1] = uni(seed, 200 + b, 4 * i + 1) * W;               // centre x
            g[i][2] = uni(seed, 200 + b, 4 * i + 2) * H;               // centre y
            g[i][3] = 20.0 + 60.0 * uni(seed, 200 + b, 4 * i + 3);     // sigma
        }
        for (int x = 0; x < W; ++x) {
            double v = 0.0;
            for (int i = 0; i < 8; ++i) {
                const double dx = x - g[i][1], dy = y - g[i][2];
                v += g[i][0] * std::exp(-(dx * dx + dy * dy) / (2.0 * g[i][3] * g[i][3]));
            }
            const uint64_t idx = ((uint64_t)b * H + y) * W + x;
            v += (2.0 * uni(seed, 5, idx) - 1.0) * 0.8 * kPi;
            double w = v - 2.0 * kPi * std::floor(v / (2.0 * kPi) + 0.5);
            if (w <= -kPi) w += 2.0 * kPi;
            if (w > kPi) w -= 2.0 * kPi;
            float f = (float)w;
            if (f > (float)kPi) f = (float)kPi;
            out[idx] = f;
        }
    });
    return CTDM_OK;
}

}  // extern "C"
