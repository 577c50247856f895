// Autotuner: for every requested volume side, times each fused volume pass (x with projection,
// y, z with data consistency) with 8 and 16 pencils per workgroup on this device, inside the
// positivity loop's own pass sequence (ctdm_time_loop; each pass reads what the previous one wrote),
// and keeps the faster choice (for y and z+dc, 8 pencils only when clearly faster) in the context
// (used by every later launch of that side, including graph captures made afterwards); then the sweep
// direction of the x and y passes (forward, or one of them walking the z planes backwards so that it
// starts on the planes the previous pass wrote last), the radix plan of the y pass and the pencils per
// workgroup of the gridding pass. An explicit call re-times every side; a reconstruction asking for
// tuning only times sides not tuned yet. The choices are also kept per (device architecture, side) for
// the rest of the process, so contexts created later start from them. The other passes keep one radix
// plan per length (fft_dev.hpp).
#include <cstdio>
#include <mutex>
#include <sstream>

#include "internal.hpp"

namespace ctdm {

namespace {
// Choices made by the autotuner in this process, per (device architecture, volume side).
std::mutex g_tuned_mu;
std::map<std::pair<std::string, int>, Tune> g_tuned;
}  // namespace

Tune& Ctx::tuning(int n)
{
    auto it = tune.find(n);
    if (it != tune.end()) return it->second;
    Tune t;
    {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        auto g = g_tuned.find({arch, n});
        if (g != g_tuned.end()) t = g->second;
    }
    return tune.emplace(n, t).first->second;
}

std::string autotune(Ctx& c, const int* sizes, int nsizes, bool force)
{
    CTDM_REQUIRE(sizes || nsizes == 0, "autotune: null sizes");
    CTDM_REQUIRE(!c.capturing, "autotune during capture");
    std::ostringstream rep;
    static const char* names[3] = {"x+proj", "y", "z+dc"};
    constexpr int kIters = 4;
    for (int i = 0; i < nsizes; ++i) {
        const int V = sizes[i];
        Tune& t = c.tuning(V);
        if (t.tuned && !force) continue;
        for (int k = 0; k < 3; ++k) {
            double m[2] = {0, 0}, bytes[3] = {0, 0, 0};
            for (int b = 0; b < 2; ++b) {  // this pass at 8 / 16 pencils, the others at their current choice
                t.pencils[k] = b ? 16 : 8;
                double ms[3];
                time_loop(c, V, kIters, ms, bytes);
                m[b] = ms[k];
            }
            t.ms[k][0] = m[0];
            t.ms[k][1] = m[1];
            // y and z+dc pencils are strided: 8 of them write 64-byte halves of 128-byte lines, which
            // costs more inside a replayed reconstruction than in this short timed sequence (measured
            // in a graph: z+dc 54.9 us at 8 pencils against 50.7 us at 16, where this timing had the two
            // within 1 %). Those passes take 8 pencils only on a clear margin; x rows are contiguous.
            const double margin = k == 0 ? 1.0 : 0.97;
            t.pencils[k] = m[0] < margin * m[1] ? 8 : 16;
            char line[200];
            std::snprintf(line, sizeof line, "V=%d %-6s B=8 %.4f ms  B=16 %.4f ms  -> B=%d (%.0f GB/s)\n", V,
                          names[k], m[0], m[1], t.pencils[k], bytes[k] / (1e6 * std::min(m[0], m[1])));
            rep << line;
        }
        // Sweep directions: the loop sequence timed with the x pass, then the y passes, walking the z
        // planes backwards (outputs are identical; only the order of the tiles changes), kept when the
        // iteration (x + 2 y + z+dc) is at least 1 % faster than with both forward.
        static const int combos[3][2] = {{0, 0}, {1, 0}, {0, 1}};
        int best = 0;
        for (int v = 0; v < 3; ++v) {
            t.rev[0] = combos[v][0];
            t.rev[1] = combos[v][1];
            double ms[3], bytes[3];
            time_loop(c, V, kIters, ms, bytes);
            t.sweep_ms[v] = ms[0] + 2.0 * ms[1] + ms[2];
            if (t.sweep_ms[v] < (best == 0 ? 0.99 : 1.0) * t.sweep_ms[best]) best = v;
        }
        t.rev[0] = combos[best][0];
        t.rev[1] = combos[best][1];
        char line[240];
        std::snprintf(line, sizeof line, "V=%d sweep  fwd %.4f ms  x-rev %.4f ms  y-rev %.4f ms per iteration -> %s\n", V,
                      t.sweep_ms[0], t.sweep_ms[1], t.sweep_ms[2], best == 0 ? "fwd" : (best == 1 ? "x-rev" : "y-rev"));
        rep << line;
        // Radix plan of the y pass: the alternative factorisation is kept when its y launches are at
        // least 1 % faster inside the loop sequence.
        for (int pv = 0; pv < 2; ++pv) {
            t.yplan = pv;
            double ms[3], bytes[3];
            time_loop(c, V, kIters, ms, bytes);
            t.yplan_ms[pv] = ms[1];
        }
        t.yplan = t.yplan_ms[1] < 0.99 * t.yplan_ms[0] ? 1 : 0;
        std::snprintf(line, sizeof line, "V=%d y plan  base %.4f ms  alternative %.4f ms -> %s\n", V, t.yplan_ms[0],
                      t.yplan_ms[1], t.yplan ? "alternative" : "base");
        rep << line;
        // Gridding pass: 8 or 16 pencils per workgroup for 60 angles (16 only when at least 3 % faster: the
        // staging of every angle's samples makes its workgroups twice as large in LDS).
        for (int b = 0; b < 2; ++b) t.grid_ms[b] = time_grid(c, V, b ? 16 : 8, 60, 5);
        t.pencils[3] = t.grid_ms[1] < 0.97 * t.grid_ms[0] ? 16 : 8;
        std::snprintf(line, sizeof line, "V=%d grid   B=8 %.4f ms  B=16 %.4f ms  -> B=%d\n", V, t.grid_ms[0],
                      t.grid_ms[1], t.pencils[3]);
        rep << line;
        t.tuned = true;
        ++c.tune_epoch;
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        g_tuned[{c.arch, V}] = t;
    }
    return rep.str();
}

}  // namespace ctdm
