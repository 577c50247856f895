// Drives libctdm through the reference-named C++ layer (include/ctdm_ref.hpp) the way the reference's
// per-frame loop and BatchRecon call do, and checks the result against the one-call batched
// pipeline (ctdm_reconstruct) with the same median step. BatchRecon runs twice: on ln A maps, and on
// amplitude maps as the reference passes them (AmpStack). The per-frame retrieval also runs on 8-bit
// frames (unsigned char, as the reference reads its BMP files) and must give the maps of float frames of
// the same values bit for bit. Prints "OK max|dn| = ..." and exits 0 on agreement.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
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
    float *dh, *db, *dw, *dla, *dun, *damp, *dhf, *dbf, *dw8, *dwf;
    unsigned char *dh8, *db8;
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
    // 8-bit frames of the same cell, and float frames holding the same integer values
    float top = 0.f;
    for (size_t i = 0; i < nh; ++i) top = std::fmax(top, std::fmax(holo[i], bg[i]));
    std::vector<unsigned char> h8(nh), b8(nh);
    std::vector<float> hf(nh), bf(nh);
    for (size_t i = 0; i < nh; ++i) {
        h8[i] = (unsigned char)std::lrint(255.f * holo[i] / top);
        b8[i] = (unsigned char)std::lrint(255.f * bg[i] / top);
        hf[i] = h8[i];
        bf[i] = b8[i];
    }
    HIPCHECK(hipMalloc(&dh8, nh));
    HIPCHECK(hipMalloc(&db8, nh));
    HIPCHECK(hipMalloc(&dhf, nh * 4));
    HIPCHECK(hipMalloc(&dbf, nh * 4));
    HIPCHECK(hipMalloc(&dw8, nm * 4));
    HIPCHECK(hipMalloc(&dwf, nm * 4));
    HIPCHECK(hipMemcpy(dh8, h8.data(), nh, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(db8, b8.data(), nh, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dhf, hf.data(), nh * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dbf, bf.data(), nh * 4, hipMemcpyHostToDevice));
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
        ctdm_crop c8{0, 0, -1, V}, cf{0, 0, -1, V};
        for (int f = 0; f < A; ++f) {
            const size_t o = (size_t)f * H * H, m = (size_t)f * V * V;
            ctdm_ref::extractQPI_AlgorithmA(c, dh8 + o, db8 + o, H, H, f + 1, c8, dw8 + m, dla + m);
            ctdm_ref::extractQPI_AlgorithmA(c, dhf + o, dbf + o, H, H, f + 1, cf, dwf + m, dla + m);
        }
        c.sync();
        if (c8.row != cf.row || c8.col != cf.col || c8.radius != cf.radius) {
            std::printf("MISMATCH: 8-bit frames located another sideband\n");
            return 1;
        }
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
    std::vector<float> w8(nm), wf(nm);
    HIPCHECK(hipMemcpy(w8.data(), dw8, nm * 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(wf.data(), dwf, nm * 4, hipMemcpyDeviceToHost));
    const bool same8 = std::memcmp(w8.data(), wf.data(), nm * 4) == 0;
    const bool ok = d <= 1e-6 && d3 <= 1e-6 && same8;
    std::printf("%s max|dn| = %.3e (ln A maps), %.3e (amplitude maps), RI %.4f..%.4f; 8-bit frames %s\n",
                ok ? "OK" : "MISMATCH", d, d3, lo, hi, same8 ? "bit-identical" : "DIFFER");
    for (void* q : {(void*)dh, (void*)db, (void*)dw, (void*)dla, (void*)dun, (void*)damp, (void*)ri1, (void*)ri2,
                    (void*)ri3, (void*)dh8, (void*)db8, (void*)dhf, (void*)dbf, (void*)dw8, (void*)dwf})
        hipFree(q);
    return ok ? 0 : 1;
}
