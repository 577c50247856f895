// Temporary: entry points of modules not yet implemented report CTDM_EUNSUPPORTED.
#include "internal.hpp"
namespace ctdm {
#define NOTYET(name) fail(CTDM_EUNSUPPORTED, name " not implemented yet")
void fit_pgd(Ctx&, float2*, const float2*, const ctdm_angle*, int, int, const ctdm_optics&, int, double, float2*, unsigned) { NOTYET("fit_pgd"); }
}
