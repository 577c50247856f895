// Temporary: entry points of modules not yet implemented report CTDM_EUNSUPPORTED.
#include "internal.hpp"
namespace ctdm {
#define NOTYET(name) fail(CTDM_EUNSUPPORTED, name " not implemented yet")



void goldstein(Ctx&, const float*, int, int, int, int8_t*, uint8_t*, int32_t*, float*) { NOTYET("goldstein"); }
void rytov_spectrum(Ctx&, const float*, const float*, int, int, float2*) { NOTYET("rytov"); }
void ewald_scatter(Ctx&, const float2*, const ctdm_angle*, int, int, int, const ctdm_optics&, float2*, float*) { NOTYET("scatter"); }
void ewald_gather(Ctx&, const float2*, const ctdm_angle*, int, int, int, const ctdm_optics&, float2*) { NOTYET("gather"); }
void normalize_measured(Ctx&, float2*, const float*, float2*, int, double) { NOTYET("normalize"); }
void fft3d(Ctx&, float2*, int, int) { NOTYET("fft3d"); }
void positivity_plain(Ctx&, float2*, const float2*, const float*, float2*, int, int, const ctdm_optics&, unsigned) { NOTYET("positivity"); }
void positivity_fused(Ctx&, float2*, const float2*, const float*, float2*, int, int, const ctdm_optics&, unsigned) { NOTYET("positivity"); }
void fit_pgd(Ctx&, float2*, const float2*, const ctdm_angle*, int, int, const ctdm_optics&, int, double, float2*, unsigned) { NOTYET("fit_pgd"); }
void reconstruct(Ctx&, ctdm_cell*, int, const ctdm_batch_opts&) { NOTYET("reconstruct"); }
std::string autotune(Ctx&, const int*, int) { NOTYET("autotune"); }
}
