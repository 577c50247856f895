// Autotuner: for every requested volume side, times each fused volume pass (x with projection,
// y, z with data consistency) with 8 and 16 pencils per workgroup on this device, inside the
// positivity loop's own pass sequence (ctdm_time_loop; each pass reads what the previous one wrote),
// and keeps the faster choice (for y and z+dc, 8 pencils only when clearly faster) in the context
// (used by every later launch of that side, including graph captures made afterwards). An explicit call re-times every side; a reconstruction asking for
// tuning only times sides not tuned yet. Radix plans are fixed per length (fft_dev.hpp); the gridding pass
// keeps 8 pencils (its LDS staging of every angle's samples limits it to two 16-pencil workgroups per CU;
// it is timed only inside a full reconstruction).
#include <cstdio>
#include <sstream>

#include "internal.hpp"

namespace ctdm {

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
        t.tuned = true;
    }
    return rep.str();
}

}  // namespace ctdm
