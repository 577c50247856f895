// Drives libctdm through the reference-named C++ layer (include/ctdm_ref.hpp) the way the reference's
// per-frame loop and BatchRecon call do, and checks the result against the one-call batched
// pipeline (ctdm_reconstruct) with the same median step. BatchRecon runs twice: on ln A maps, and on
// amplitude maps as the reference passes them (AmpStack). Prints "OK max|dn| = ..." and exits 0 on
// agreement.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "ctdm_ref.hpp"

#define HIPCHECK(x)                                                              \
    do {                                                                         \
        if ((x) != hipSuccess) {                                                 \
            std::printf("HIP error %s at line %d\n", #x, __LINE__);              \
            return 2;                                                            \
        }                                                                        \
    } while (0)

int main()
{
    const int A = 8, H = 256, V = 64, iters = 3;
    ctdm_synth_params p;
    ctdm_synth_default_params(&p);
    p.H = H;
    std::vector<float> holo((size_t)A * H * H), bg((size_t)A * H * H);
    std::vector<ctdm_angle> ang(A);
    if (ctdm_synth_cell(7, A, &p, holo.data(), bg.data(), ang.data()) != CTDM_OK) return 2;
    float *dh, *db, *dw, *dla, *dun, *damp;
    ctdm_complex *ri1, *ri2, *ri3;
    const size_t nh = (size_t)A * H * H, nm = (size_t)A * V * V, nv = (size_t)V * V * V;
    HIPCHECK(hipMalloc(&dh, nh * 4));
    HIPCHECK(hipMalloc(&db, nh * 4));
    HIPCHECK(hipMalloc(&dw, nm * 4));
    HIPCHECK(hipMalloc(&dla, nm * 4));
    HIPCHECK(hipMalloc(&dun, nm * 4));
    HIPCHECK(hipMalloc(&damp, nm * 4));
    HIPCHECK(hipMalloc(&ri1, nv * 8));
    HIPCHECK(hipMalloc(&ri2, nv * 8));
    HIPCHECK(hipMalloc(&ri3, nv * 8));
    HIPCHECK(hipMemcpy(dh, holo.data(), nh * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(db, bg.data(), nh * 4, hipMemcpyHostToDevice));
    try {
        ctdm_ref::Context c;
        ctdm_crop crop{0, 0, -1, V};
        for (int f = 0; f < A; ++f) {  // per-frame loop (HilbertTransform.cu:785-1165)
            ctdm_ref::extractQPI_AlgorithmA(c, dh + (size_t)f * H * H, db + (size_t)f * H * H, H, H, f + 1, crop,
                                            dw + (size_t)f * V * V, dla + (size_t)f * V * V);
            ctdm_ref::DCT_UWLS_Unwrapped(c, dun + (size_t)f * V * V, dw + (size_t)f * V * V, V, V);
        }
        std::vector<float> ax(A), ay(A);
        for (int i = 0; i < A; ++i) {
            ax[i] = (float)ang[i].ang_x;
            ay[i] = (float)ang[i].ang_y;
        }
        std::vector<ctdm_angle> angf(A);
        for (int i = 0; i < A; ++i) angf[i] = ctdm_angle{ax[i], ay[i]};
        const ctdm_optics o{532e-9, 1.337, 0.1e-6 * H / V};
        ctdm_ref::BatchRecon(c, dun, dla, ax.data(), ay.data(), A, V, o, iters, ri1, CTDM_MAPS_MEDIAN);
        c.sync();
        std::vector<float> h(nm);  // amplitude stack as the reference hands it over
        HIPCHECK(hipMemcpy(h.data(), dla, nm * 4, hipMemcpyDeviceToHost));
        for (float& v : h) v = std::exp(v);
        HIPCHECK(hipMemcpy(damp, h.data(), nm * 4, hipMemcpyHostToDevice));
        ctdm_ref::BatchRecon(c, dun, damp, ax.data(), ay.data(), A, V, o, iters, ri3);
        ctdm_cell cell{};
        cell.holo = dh;
        cell.bg = db;
        cell.angles = angf.data();
        cell.A = A;
        cell.H = H;
        cell.W = H;
        cell.V = V;
        cell.iters = iters;
        cell.optics = o;
        cell.crop = crop;
        cell.ri = ri2;
        cell.recon_flags = CTDM_MAPS_MEDIAN;
        ctdm_batch_opts opts{};
        opts.mode = 1;
        c.check(ctdm_reconstruct(c.get(), &cell, 1, &opts));
        c.sync();
    } catch (const ctdm_ref::Error& e) {
        std::printf("ctdm error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    std::vector<ctdm_complex> a(nv), b(nv), e(nv);
    HIPCHECK(hipMemcpy(a.data(), ri1, nv * 8, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(b.data(), ri2, nv * 8, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(e.data(), ri3, nv * 8, hipMemcpyDeviceToHost));
    double d = 0, d3 = 0, lo = 1e9, hi = 0;
    for (size_t i = 0; i < nv; ++i) {
        d = std::fmax(d, std::fabs(a[i].x - b[i].x) + std::fabs(a[i].y - b[i].y));
        d3 = std::fmax(d3, std::fabs(e[i].x - b[i].x) + std::fabs(e[i].y - b[i].y));
        lo = std::fmin(lo, a[i].x);
        hi = std::fmax(hi, a[i].x);
    }
    const bool ok = d <= 1e-6 && d3 <= 1e-6;
    std::printf("%s max|dn| = %.3e (ln A maps), %.3e (amplitude maps), RI %.4f..%.4f\n", ok ? "OK" : "MISMATCH", d,
                d3, lo, hi);
    for (void* q : {(void*)dh, (void*)db, (void*)dw, (void*)dla, (void*)dun, (void*)damp, (void*)ri1, (void*)ri2,
                    (void*)ri3})
        hipFree(q);
    return ok ? 0 : 1;
}
