// Autotuner: for every requested volume side, times each fused volume pass (x with projection,
// y, z with data consistency) with 8 and 16 pencils per workgroup on this device and keeps the
// faster choice in the context (used by every later launch of that side, including graph
// captures made afterwards). An explicit call re-times every side; a reconstruction asking for
// tuning only times sides not tuned yet. Radix plans are fixed per length (fft_dev.hpp); the gridding pass
// keeps 16 pencils (it is timed only inside a full reconstruction).
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
    for (int i = 0; i < nsizes; ++i) {
        const int V = sizes[i];
        Tune& t = c.tuning(V);
        if (t.tuned && !force) continue;
        for (int k = 0; k < 3; ++k) {
            double bytes = 0;
            const double m8 = time_pass(c, k, V, 8, 10, &bytes);
            const double m16 = time_pass(c, k, V, 16, 10, &bytes);
            t.ms[k][0] = m8;
            t.ms[k][1] = m16;
            t.pencils[k] = m8 < m16 ? 8 : 16;
            char line[200];
            std::snprintf(line, sizeof line, "V=%d %-6s B=8 %.4f ms  B=16 %.4f ms  -> B=%d (%.0f GB/s)\n", V,
                          names[k], m8, m16, t.pencils[k], bytes / (1e6 * std::min(m8, m16)));
            rep << line;
        }
        t.tuned = true;
    }
    return rep.str();
}

}  // namespace ctdm
