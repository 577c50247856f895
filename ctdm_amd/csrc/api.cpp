// C ABI of libctdm: argument checking, status conversion and the context object. The compute lives
// in the k_*.hip modules; this file never launches kernels itself.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "internal.hpp"

struct ctdm_ctx {
    ctdm::Ctx c;
};

namespace ctdm {

const float2* Ctx::tw(int n)
{
    auto it = twiddles.find(n);
    if (it != twiddles.end()) return it->second;
    CTDM_REQUIRE(!capturing, "twiddle table requested during graph capture");
    std::vector<float2> h(n);
    for (int m = 0; m < n; ++m) {
        const double a = -2.0 * 3.14159265358979323846 * (double)m / (double)n;
        h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    float2* d = nullptr;
    CTDM_HIP(hipMalloc(&d, sizeof(float2) * n));
    CTDM_HIP(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    twiddles[n] = d;
    return d;
}

void* Ctx::buf(const std::string& name, size_t bytes, bool shared)
{
    DevBuf& b = scratch[lane && !shared ? name + "#lane1" : name];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    CTDM_REQUIRE(!capturing, "scratch buffer '" + name + "' grown during graph capture");
    if (b.ptr) {
        CTDM_HIP(hipStreamSynchronize(stream));
        if (referenced(b.ptr))
            retired.push_back(b.ptr);
        else
            release(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    CTDM_HIP(hipMalloc(&b.ptr, bytes ? bytes : 1));
    b.bytes = bytes;
    return b.ptr;
}

bool Ctx::referenced(const void* p) const
{
    for (const auto& kv : graphs)
        if (std::binary_search(kv.second.ptrs.begin(), kv.second.ptrs.end(), (void*)p)) return true;
    return false;
}

void Ctx::trim_graphs(size_t keep)
{
    if (graphs.size() <= keep) return;
    CTDM_REQUIRE(!capturing, "graph cache trimmed during capture");
    CTDM_HIP(hipStreamSynchronize(stream));  // no evicted graph may still be running
    while (graphs.size() > keep) {
        auto victim = graphs.begin();
        for (auto it = graphs.begin(); it != graphs.end(); ++it)
            if (it->second.last_use < victim->second.last_use) victim = it;
        if (victim->second.exec) CTDM_HIP(hipGraphExecDestroy(victim->second.exec));
        if (victim->second.graph) CTDM_HIP(hipGraphDestroy(victim->second.graph));
        graphs.erase(victim);
    }
    std::vector<void*> still;
    for (void* p : retired) {
        if (referenced(p))
            still.push_back(p);
        else
            release(p);
    }
    retired.swap(still);
}

void Ctx::release(void* p)
{
    table_contents.erase(p);
    CTDM_HIP(hipFree(p));
}

void Ctx::ensure_aux()
{
    if (aux) return;
    CTDM_REQUIRE(!capturing, "auxiliary stream created during graph capture");
    CTDM_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    for (hipEvent_t& e : lane_ev) CTDM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

unsigned long long* Ctx::host_words(int n)
{
    CTDM_REQUIRE(n > 0, "host_words: empty request");
    if (n > host_words_n_) {
        if (host_words_) CTDM_HIP(hipHostFree(host_words_));
        host_words_ = nullptr;
        host_words_n_ = 0;
        const int cap = std::max(n, 64);
        CTDM_HIP(hipHostMalloc((void**)&host_words_, cap * sizeof(unsigned long long), hipHostMallocDefault));
        host_words_n_ = cap;
    }
    return host_words_;
}

void Ctx::spin_sync()
{
    for (;;) {
        const hipError_t e = hipStreamQuery(stream);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) CTDM_HIP(e);
    }
}

void Ctx::spin_event(hipEvent_t e)
{
    for (;;) {
        const hipError_t q = hipEventQuery(e);
        if (q == hipSuccess) return;
        if (q != hipErrorNotReady) CTDM_HIP(q);
    }
}

}  // namespace ctdm

namespace {

template <class F>
ctdm_status guarded(ctdm_ctx* ctx, F&& f)
{
    if (!ctx) return CTDM_EINVAL;
    try {
        if (hipSetDevice(ctx->c.device) != hipSuccess) ctdm::fail(CTDM_EDEVICE, "hipSetDevice failed");
        f(ctx->c);
        return CTDM_OK;
    } catch (const ctdm::Error& e) {
        ctx->c.last_error = e.msg;
        return e.status;
    } catch (const std::bad_alloc&) {
        ctx->c.last_error = "host allocation failed";
        return CTDM_ENOMEM;
    } catch (...) {
        ctx->c.last_error = "unknown error";
        return CTDM_EDEVICE;
    }
}

void check_optics(const ctdm_optics* o)
{
    CTDM_REQUIRE(o, "optics: null pointer");
    CTDM_REQUIRE(o->wavelength > 0 && o->n_med > 0 && o->dx > 0, "optics: non-positive parameter");
}

}  // namespace

extern "C" {

const char* ctdm_version(void) { return "ctdm-mi355x 0.1 (gfx950)"; }

ctdm_status ctdm_create(ctdm_ctx** out, const ctdm_config* cfg, ctdm_stream stream)
{
    if (!out) return CTDM_EINVAL;
    *out = nullptr;
    ctdm_ctx* ctx = new (std::nothrow) ctdm_ctx();
    if (!ctx) return CTDM_ENOMEM;
    ctx->c.device = cfg ? cfg->device : 0;
    ctx->c.stream = (hipStream_t)stream;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || ctx->c.device < 0 || ctx->c.device >= n) {
        delete ctx;
        return CTDM_EDEVICE;
    }
    if (hipSetDevice(ctx->c.device) != hipSuccess) {
        delete ctx;
        return CTDM_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->c.device) == hipSuccess) {
        ctx->c.max_lds = (int)prop.sharedMemPerBlock;
        ctx->c.arch = prop.gcnArchName;
        ctx->c.cus = prop.multiProcessorCount;
    }
    const ctdm_status st = guarded(ctx, [](ctdm::Ctx& c) { ctdm::noalloc_setup(c); });
    if (st != CTDM_OK) {
        ctdm_destroy(ctx);
        return st;
    }
    *out = ctx;
    return CTDM_OK;
}

ctdm_status ctdm_destroy(ctdm_ctx* ctx)
{
    if (!ctx) return CTDM_EINVAL;
    hipSetDevice(ctx->c.device);
    hipStreamSynchronize(ctx->c.stream);
    for (auto& kv : ctx->c.graphs) {
        if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) hipGraphDestroy(kv.second.graph);
    }
    for (auto& kv : ctx->c.twiddles) hipFree(kv.second);
    for (auto& kv : ctx->c.scratch)
        if (kv.second.ptr) hipFree(kv.second.ptr);
    for (void* p : ctx->c.retired) hipFree(p);
    if (ctx->c.host_words_) hipHostFree(ctx->c.host_words_);
    if (ctx->c.aux) {
        hipStreamSynchronize(ctx->c.aux);
        hipStreamDestroy(ctx->c.aux);
    }
    for (hipEvent_t e : ctx->c.lane_ev)
        if (e) hipEventDestroy(e);
    if (ctx->c.wait_ev) hipEventDestroy(ctx->c.wait_ev);
    delete ctx;
    return CTDM_OK;
}

ctdm_status ctdm_set_stream(ctdm_ctx* ctx, ctdm_stream stream)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        CTDM_HIP(hipStreamSynchronize(c.stream));
        c.stream = (hipStream_t)stream;
    });
}

ctdm_status ctdm_sync(ctdm_ctx* ctx)
{
    return guarded(ctx, [&](ctdm::Ctx& c) { CTDM_HIP(hipStreamSynchronize(c.stream)); });
}

const char* ctdm_last_error(const ctdm_ctx* ctx)
{
    if (!ctx) return "null context";
    return ctx->c.last_error.c_str();
}

namespace {

// Pixel format of an image argument; an unknown value is a caller error.
ctdm::Pixels pixels(const void* p, ctdm_pixel px)
{
    CTDM_REQUIRE(px == CTDM_PIXEL_F32 || px == CTDM_PIXEL_U8 || px == CTDM_PIXEL_U16, "unknown pixel format");
    return ctdm::Pixels{p, (ctdm::Px)px};
}

}  // namespace

ctdm_status ctdm_find_sideband_px(ctdm_ctx* ctx, const void* bg, ctdm_pixel px, int H, int W, int M, int row_start,
                                  ctdm_crop* out)
{
    return guarded(ctx, [&](ctdm::Ctx& c) { ctdm::find_sideband(c, pixels(bg, px), H, W, M, row_start, out); });
}

ctdm_status ctdm_find_sideband(ctdm_ctx* ctx, const float* bg, int H, int W, int M, int row_start,
                               ctdm_crop* out)
{
    return ctdm_find_sideband_px(ctx, bg, CTDM_PIXEL_F32, H, W, M, row_start, out);
}

ctdm_status ctdm_retrieve_px(ctdm_ctx* ctx, const void* holo, const void* bg, ctdm_pixel px, int batch, int H, int W,
                             const ctdm_crop* crop, float* wrapped, float* logamp, ctdm_complex* field)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        CTDM_REQUIRE(crop, "retrieve: null crop");
        ctdm::retrieve_batched(c, pixels(holo, px), pixels(bg, px), batch, H, W, *crop, wrapped, logamp,
                               (float2*)field);
    });
}

ctdm_status ctdm_retrieve(ctdm_ctx* ctx, const float* holo, const float* bg, int batch, int H, int W,
                          const ctdm_crop* crop, float* wrapped, float* logamp, ctdm_complex* field)
{
    return ctdm_retrieve_px(ctx, holo, bg, CTDM_PIXEL_F32, batch, H, W, crop, wrapped, logamp, field);
}

ctdm_status ctdm_find_sideband_rows_px(ctdm_ctx* ctx, const void* bg, ctdm_pixel px, int H, int W, int* center)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        CTDM_REQUIRE(center, "find_sideband_rows: null output");
        *center = ctdm::find_sideband_rows(c, pixels(bg, px), H, W);
    });
}

ctdm_status ctdm_find_sideband_rows(ctdm_ctx* ctx, const float* bg, int H, int W, int* center)
{
    return ctdm_find_sideband_rows_px(ctx, bg, CTDM_PIXEL_F32, H, W, center);
}

ctdm_status ctdm_retrieve_rows_px(ctdm_ctx* ctx, const void* holo, const void* bg, ctdm_pixel px, int batch, int H,
                                  int W, int center, float* wrapped, float* logamp)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        ctdm::retrieve_rows(c, pixels(holo, px), pixels(bg, px), batch, H, W, center, wrapped, logamp);
    });
}

ctdm_status ctdm_retrieve_rows(ctdm_ctx* ctx, const float* holo, const float* bg, int batch, int H, int W, int center,
                               float* wrapped, float* logamp)
{
    return ctdm_retrieve_rows_px(ctx, holo, bg, CTDM_PIXEL_F32, batch, H, W, center, wrapped, logamp);
}

ctdm_status ctdm_angle_from_image_px(ctdm_ctx* ctx, const void* img, ctdm_pixel px, int batch, int H, int W,
                                     double pixel_um, double mag, double wavelength_nm, double n_med, double* angles)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        CTDM_REQUIRE(batch >= 0, "angle_from_image: negative batch");
        const ctdm::Pixels im = pixels(img, px);
        if (batch > 0) ctdm::angle_from_image(c, im, batch, H, W, pixel_um, mag, wavelength_nm, n_med, angles);
    });
}

ctdm_status ctdm_angle_from_image(ctdm_ctx* ctx, const float* img, int batch, int H, int W, double pixel_um,
                                  double mag, double wavelength_nm, double n_med, double* angles)
{
    return ctdm_angle_from_image_px(ctx, img, CTDM_PIXEL_F32, batch, H, W, pixel_um, mag, wavelength_nm, n_med,
                                    angles);
}

ctdm_status ctdm_unwrap_dct(ctdm_ctx* ctx, const float* wrapped, float* unwrapped, int batch, int M,
                            unsigned flags)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        if (flags & CTDM_EXEC_PLAIN) {
            for (int b = 0; b < batch; ++b)
                ctdm::unwrap_dct_plain_one(c, wrapped + (size_t)b * M * M,
                                           unwrapped + (size_t)b * M * M, M, flags & 0xffu);
        } else {
            ctdm::unwrap_dct_batched(c, wrapped, unwrapped, batch, M, flags);
        }
    });
}

ctdm_status ctdm_unwrap_dct_noalloc(ctdm_ctx* ctx, const float* wrapped, float* unwrapped, int batch, int M,
                                    unsigned flags)
{
    return guarded(ctx, [&](ctdm::Ctx& c) { ctdm::unwrap_dct_noalloc(c, wrapped, unwrapped, batch, M, flags); });
}

ctdm_status ctdm_unwrap_dct_adjoint(ctdm_ctx* ctx, const float* wrapped, const float* g_out,
                                    float* g_in, int batch, int M, unsigned flags)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        ctdm::unwrap_dct_adjoint(c, wrapped, g_out, g_in, batch, M, flags);
    });
}

ctdm_status ctdm_goldstein(ctdm_ctx* ctx, const float* wrapped, int batch, int H, int W,
                           int8_t* residues, uint8_t* cuts, int32_t* k2pi, float* unwrapped)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        ctdm::goldstein(c, wrapped, batch, H, W, residues, cuts, k2pi, unwrapped);
    });
}

ctdm_status ctdm_rytov_spectrum(ctdm_ctx* ctx, const float* unwrapped, const float* logamp, int A,
                                int M, ctdm_complex* spec)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        ctdm::rytov_spectrum(c, unwrapped, logamp, A, M, (float2*)spec);
    });
}

ctdm_status ctdm_ewald_scatter(ctdm_ctx* ctx, const ctdm_complex* spec, const ctdm_angle* angles,
                               int A, int M, int V, const ctdm_optics* optics, ctdm_complex* F,
                               float* hits)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        check_optics(optics);
        ctdm::ewald_scatter(c, (const float2*)spec, angles, A, M, V, *optics, (float2*)F, hits);
    });
}

ctdm_status ctdm_ewald_gather(ctdm_ctx* ctx, const ctdm_complex* F, const ctdm_angle* angles, int A,
                              int M, int V, const ctdm_optics* optics, ctdm_complex* spec)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        check_optics(optics);
        ctdm::ewald_gather(c, (const float2*)F, angles, A, M, V, *optics, (float2*)spec);
    });
}

ctdm_status ctdm_normalize_measured(ctdm_ctx* ctx, ctdm_complex* F, const float* hits,
                                    ctdm_complex* F_meas, int V, double dx)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        ctdm::normalize_measured(c, (float2*)F, hits, (float2*)F_meas, V, dx);
    });
}

ctdm_status ctdm_fft3d(ctdm_ctx* ctx, ctdm_complex* vol, int V, int dir)
{
    return guarded(ctx, [&](ctdm::Ctx& c) { ctdm::fft3d(c, (float2*)vol, V, dir); });
}

ctdm_status ctdm_positivity(ctdm_ctx* ctx, ctdm_complex* f, const ctdm_complex* F_meas,
                            const float* hits, ctdm_complex* n, int V, int iters,
                            const ctdm_optics* optics, unsigned flags)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        check_optics(optics);
        if (flags & CTDM_EXEC_PLAIN)
            ctdm::positivity_plain(c, (float2*)f, (const float2*)F_meas, hits, (float2*)n, V, iters,
                                   *optics, flags);
        else
            ctdm::positivity_fused(c, (float2*)f, (const float2*)F_meas, hits, (float2*)n, V, iters,
                                   *optics, flags);
    });
}

ctdm_status ctdm_median3x3(ctdm_ctx* ctx, const float* in, float* out, int batch, int H, int W)
{
    return guarded(ctx, [&](ctdm::Ctx& c) { ctdm::median3x3(c, in, out, batch, H, W); });
}

ctdm_status ctdm_check_maps(ctdm_ctx* ctx, float* phase, float* amp, size_t n)
{
    return guarded(ctx, [&](ctdm::Ctx& c) { ctdm::check_maps(c, phase, amp, (long long)n); });
}

ctdm_status ctdm_recon_from_maps(ctdm_ctx* ctx, const float* unwrapped, const float* logamp,
                                 const ctdm_angle* angles, int A, int V, const ctdm_optics* optics, int iters,
                                 unsigned recon_flags, ctdm_complex* ri)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        check_optics(optics);
        ctdm::recon_from_maps(c, unwrapped, logamp, angles, A, V, *optics, iters, recon_flags, (float2*)ri);
    });
}

ctdm_status ctdm_reconstruct(ctdm_ctx* ctx, ctdm_cell* cells, int ncells, const ctdm_batch_opts* opts)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        CTDM_REQUIRE(cells || ncells == 0, "reconstruct: null cells");
        ctdm_batch_opts o{};
        if (opts) o = *opts;
        ctdm::reconstruct(c, cells, ncells, o);
    });
}

ctdm_status ctdm_autotune(ctdm_ctx* ctx, const int* sizes, int nsizes, char* report, size_t len)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        const std::string r = ctdm::autotune(c, sizes, nsizes, true);
        if (report && len) {
            std::strncpy(report, r.c_str(), len - 1);
            report[len - 1] = 0;
        }
    });
}

ctdm_status ctdm_get_tuning(ctdm_ctx* ctx, int V, int* values, int n)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        CTDM_REQUIRE(values && n == 7, "get_tuning: 7 values");
        const ctdm::Tune& t = c.tuning(V);
        for (int k = 0; k < 4; ++k) values[k] = t.pencils[k];
        values[4] = t.rev[0];
        values[5] = t.rev[1];
        values[6] = t.yplan;
    });
}

ctdm_status ctdm_set_tuning(ctdm_ctx* ctx, int V, const int* values, int n)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        CTDM_REQUIRE(values && n == 7, "set_tuning: 7 values");
        CTDM_REQUIRE(V == 32 || V == 64 || V == 128 || V == 192 || V == 256, "set_tuning: volume side has no kernel");
        for (int k = 0; k < 4; ++k) CTDM_REQUIRE(values[k] == 8 || values[k] == 16, "set_tuning: pencils must be 8 or 16");
        CTDM_REQUIRE((values[4] | values[5] | values[6]) >> 1 == 0, "set_tuning: sweeps and plan must be 0 or 1");
        ctdm::Tune& t = c.tuning(V);
        for (int k = 0; k < 4; ++k) t.pencils[k] = values[k];
        t.rev[0] = values[4];
        t.rev[1] = values[5];
        t.yplan = values[6];
        t.tuned = true;
        ++c.tune_epoch;
    });
}

ctdm_status ctdm_time_pass(ctdm_ctx* ctx, int pass, int V, int pencils, int reps, double* avg_ms,
                           double* bytes)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        CTDM_REQUIRE(avg_ms && bytes, "time_pass: null output");
        if (pencils <= 0) pencils = c.tuning(V).pencils[pass < 0 || pass > 3 ? 0 : pass];
        *avg_ms = ctdm::time_pass(c, pass, V, pencils, reps, bytes);
    });
}

ctdm_status ctdm_time_loop(ctdm_ctx* ctx, int V, int iters, double* avg_ms, double* bytes)
{
    return guarded(ctx, [&](ctdm::Ctx& c) { ctdm::time_loop(c, V, iters, avg_ms, bytes); });
}

ctdm_status ctdm_fit_pgd(ctdm_ctx* ctx, ctdm_complex* f, const ctdm_complex* spec,
                         const ctdm_angle* angles, int A, int V, const ctdm_optics* optics, int iters,
                         double step, ctdm_complex* n, unsigned flags)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        check_optics(optics);
        ctdm::fit_pgd(c, (float2*)f, (const float2*)spec, angles, A, V, *optics, iters, step,
                      (float2*)n, flags);
    });
}

ctdm_status ctdm_fit_pgd_trace(ctdm_ctx* ctx, ctdm_complex* f, const ctdm_complex* spec,
                               const ctdm_angle* angles, int A, int V, const ctdm_optics* optics, int iters,
                               double step, ctdm_complex* n, unsigned flags, double* objective)
{
    return guarded(ctx, [&](ctdm::Ctx& c) {
        check_optics(optics);
        CTDM_REQUIRE(objective || iters == 0, "fit_pgd_trace: null objective array");
        ctdm::fit_pgd(c, (float2*)f, (const float2*)spec, angles, A, V, *optics, iters, step,
                      (float2*)n, flags, objective);
    });
}

}  // extern "C"
