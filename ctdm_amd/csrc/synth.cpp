// Seeded synthetic inputs for the cTDM path (host C++). The reference reads BMP holograms from
// disk (HilbertTransform.cu:2680-2731); these generators produce the same kind of data in memory.
//
// Random numbers are counter-based (SplitMix64 of seed, stream and index), so every pixel is
// reproducible independently of threading.
//
// Models
//   hologram   I = A^2 + 1 + 2 A cos(2 pi (fx x + fy y) + phi) + N(0, sigma^2); background A = 1,
//              phi = 0 with independent noise.
//   cell       phase and log-amplitude of each illumination angle are straight-ray projections
//              (the geometric-optics limit of the first-order Rytov phase) of a phantom made of an
//              ellipsoid (n_cell + i kappa) holding spherical nucleoli (n_nucleoli):
//                 phi_a(x, y) = k0 * integral (Re n - n_med) dt,  ln A_a = -k0 * integral Im n dt
//              along the illumination direction d_a through (x, y, 0). The angles lie on a cone
//              n_med sin(theta) = NA_ill with azimuth 2 pi a / A; ang_y = asin(d_x),
//              ang_x = asin(d_y) so that the Ewald mapping (f_x0 = f_m0 sin(ang_y),
//              f_y0 = f_m0 sin(ang_x)) points along d_a.
//   wrapped    sum of 8 Gaussians (|peak| <= 40 rad, sigma 20-80 px) + U(-0.8 pi, 0.8 pi), wrapped
//              to (-pi, pi].
#include <algorithm>
#include <cmath>
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
    if (!p || !holo || !bg || !angles || A <= 0 || p->H < 16) return CTDM_EINVAL;
    const int H = p->H;
    const double um = 1e6;  // work in micrometres
    const double pix = p->pixel * um, k0 = 2.0 * kPi / (p->wavelength * um);
    Ellipsoid cell{{0, 0, 0}, {p->semi_axes[0] * um, p->semi_axes[1] * um, p->semi_axes[2] * um}};
    std::vector<Ellipsoid> nuc(p->n_nuc);
    for (int i = 0; i < p->n_nuc; ++i) {
        // uniform in the inner half of the cell (keeps nucleoli inside the ellipsoid)
        double v[3];
        do {
            for (int a = 0; a < 3; ++a) v[a] = 2.0 * uni(seed, 100 + i, a) - 1.0;
            seed += 0x51ed27;
        } while (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > 1.0);
        for (int a = 0; a < 3; ++a) {
            nuc[i].c[a] = 0.5 * v[a] * cell.s[a];
            nuc[i].s[a] = p->r_nuc * um;
        }
    }
    const double sin_t = p->na_ill / p->n_med;
    const double cos_t = std::sqrt(std::max(0.0, 1.0 - sin_t * sin_t));
    for (int a = 0; a < A; ++a) {
        const double az = 2.0 * kPi * a / A;
        const double dx = sin_t * std::cos(az), dy = sin_t * std::sin(az);
        angles[a].ang_y = std::asin(dx);
        angles[a].ang_x = std::asin(dy);
    }
    const double dn_cell = p->n_cell - p->n_med, dn_nuc = p->n_nucleoli - p->n_cell;
    parallel_rows(A * H, [&](int ay) {
        const int a = ay / H, y = ay % H;
        const double az = 2.0 * kPi * a / A;
        const double d[3] = {sin_t * std::cos(az), sin_t * std::sin(az), cos_t};
        float* hp = holo + (size_t)a * H * H + (size_t)y * H;
        float* bp = bg + (size_t)a * H * H + (size_t)y * H;
        for (int x = 0; x < H; ++x) {
            const double p0[3] = {(x - H / 2) * pix, (y - H / 2) * pix, 0.0};
            double phi = 0.0, la = 0.0;
            const double lc = chord(cell, p0, d);
            if (lc > 0) {
                phi = k0 * dn_cell * lc;
                la = -k0 * p->kappa_cell * lc;
                for (const auto& e : nuc) phi += k0 * dn_nuc * chord(e, p0, d);
            }
            const double Aamp = std::exp(la);
            const double carrier = 2.0 * kPi * (p->carrier[0] * x + p->carrier[1] * y);
            const uint64_t idx = ((uint64_t)a * H + y) * H + x;
            // sample field E = A exp(i phi) (+ complex noise); I = |E|^2 + 1 + 2 Re(E exp(i carrier))
            double er = Aamp * std::cos(phi), ei = Aamp * std::sin(phi);
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
