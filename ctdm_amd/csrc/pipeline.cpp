// Cell pipelines: hologram pairs -> RI volume.
//
// plain   (mode 0): the reference's structure rebuilt on this library's kernels — per angle,
//                   full 2-D transforms, separate shift/crop/scale kernels, the 14-kernel DCT
//                   unwrapper, per-angle Rytov FFT and scatter, unfused positivity loop (10 volume
//                   passes per iteration), host-synchronous between stages (the reference's
//                   cudaMemcpy / cudaThreadSynchronize points, HilbertTransform.cu:785-1165,
//                   1825-2047, FFT.cu:655-736).
// batched (mode 1): all angles of a cell per launch, fused passes (3 retrieval, 3 unwrap,
//                   2 Rytov, 3 + 4*iters volume kernels), optionally captured once into a hipGraph
//                   and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace ctdm {

void scatter_plain_one(Ctx& c, const float2* u_sp, const AngleRow& r, int V, const EwaldGeom& g, float2* F,
                       float* hits);
void zero_volume(Ctx& c, float2* F, float* hits, int V);

// Pixel format of a cell's holograms and backgrounds: 8- or 16-bit integers with CTDM_HOLO_U8 or
// CTDM_HOLO_U16, float32 otherwise.
static Px cell_px(const ctdm_cell& cl)
{
    const unsigned f = cl.recon_flags & (CTDM_HOLO_U8 | CTDM_HOLO_U16);
    return f == CTDM_HOLO_U8 ? Px::U8 : (f == CTDM_HOLO_U16 ? Px::U16 : Px::F32);
}
static Pixels holo_of(const ctdm_cell& cl) { return Pixels{cl.holo, cell_px(cl)}; }
static Pixels bg_of(const ctdm_cell& cl) { return Pixels{cl.bg, cell_px(cl)}; }

static void validate_cell(const ctdm_cell& cl)
{
    CTDM_REQUIRE(cl.holo && cl.bg && cl.angles && cl.ri, "cell: null pointer");
    CTDM_REQUIRE((cl.recon_flags & (CTDM_HOLO_U8 | CTDM_HOLO_U16)) != (CTDM_HOLO_U8 | CTDM_HOLO_U16),
                 "cell: CTDM_HOLO_U8 and CTDM_HOLO_U16 both set");
    CTDM_REQUIRE(cl.A > 0 && cl.H > 0 && cl.W > 0, "cell: empty");
    CTDM_REQUIRE(cl.V == 32 || cl.V == 64 || cl.V == 128 || cl.V == 192 || cl.V == 256,
                 "cell: volume side has no kernel");
    CTDM_REQUIRE(cl.iters >= 0, "cell: negative iteration count");
    CTDM_REQUIRE(cl.crop.radius < 0 || cl.crop.M == cl.V, "cell: crop side must equal the volume side");
    if (cl.recon_flags & CTDM_RETRIEVE_ROWS)
        CTDM_REQUIRE(cl.H == cl.W && cl.V * 4 == cl.W, "cell: row-wise retrieval needs square holograms and V = W / 4");
    CTDM_REQUIRE(cl.optics.wavelength > 0 && cl.optics.n_med > 0 && cl.optics.dx > 0, "cell: bad optics");
}

// Device crop of batched cell i of a call. It is written on the context stream before the cells' work
// each call (by the sideband search or from the given crop), so the captured sequence does not depend
// on the carrier position. Its size never changes, so every graph sees the same slot.
static CropDev* crop_slot(Ctx& c, int i)
{
    return (CropDev*)c.buf("bt.crop." + std::to_string(i), sizeof(CropDev), true);
}

// Device angle-row table of batched cell i, sized for A angles. Like the crop slot it is written before
// the cells' work each call, eager or replayed, so a graph follows new angle values in the same buffers.
static double* rows_slot(Ctx& c, int i, int A)
{
    return (double*)c.buf("bt.rows." + std::to_string(i), sizeof(double) * 3 * (size_t)std::max(A, 1), true);
}

static ctdm_crop cell_crop(Ctx& c, const ctdm_cell& cl)
{
    if (cl.crop.radius >= 0) return cl.crop;
    if (cl.recon_flags & CTDM_RETRIEVE_ROWS) {  // sideband column located on frame 1 (HilbertTransform.cu:1634)
        const int center = find_sideband_rows(c, bg_of(cl), cl.H, cl.W);
        return ctdm_crop{0, center, cl.W / 8, cl.V};
    }
    ctdm_crop cr{};
    find_sideband(c, bg_of(cl), cl.H, cl.W, cl.V, 0, &cr);  // frame 1 of the reference (HilbertTransform.cu:1572)
    return cr;
}

__global__ void k_wrap_log(const float* __restrict__ phase, const float* __restrict__ amp, float* __restrict__ w,
                           float* __restrict__ la, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (w) {
        const float p = phase[i];
        w[i] = p - 6.283185307179586f * rintf(p / 6.283185307179586f);
    }
    if (la) {
        const float l = logf(amp[i]);
        la[i] = isfinite(l) ? l : 0.f;
    }
}

// amp = exp(la), elementwise (la and amp may be the same array).
__global__ void k_exp(const float* la, float* amp, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) amp[i] = expf(la[i]);
}

static void run_plain(Ctx& c, ctdm_cell& cl)
{
    const int A = cl.A, H = cl.H, W = cl.W, V = cl.V;
    const size_t img = (size_t)V * V;
    cl.crop = cell_crop(c, cl);
    float* phase = (float*)c.buf("pl.phase", sizeof(float) * A * img);
    float* amp = (float*)c.buf("pl.amp", sizeof(float) * A * img);
    float* unw = cl.unwrapped_opt ? cl.unwrapped_opt : (float*)c.buf("pl.unw", sizeof(float) * A * img);
    const bool by_rows = cl.recon_flags & CTDM_RETRIEVE_ROWS;
    const bool fast = cl.unwrap_flags & CTDM_UNWRAP_REFCOMPAT_FAST;
    for (int a = 0; a < A; ++a) {
        const Pixels sp = holo_of(cl).offset((size_t)a * H * W), bk = bg_of(cl).offset((size_t)a * H * W);
        if (by_rows) {  // one angle per call; the map is ln A, turned into A for the reference's copy kernel
            // the phase map holds calcWrapPhase's value, as after retrieve_plain_one (FastUnwrapping's
            // min-max rescaling depends on it; every other unwrapping is the same for the wrapped map)
            float* wr1 = fast ? (float*)c.buf("pl.wr1", sizeof(float) * img) : phase + a * img;
            retrieve_rows(c, sp, bk, 1, H, W, cl.crop.col, wr1, amp + a * img, fast ? phase + a * img : nullptr);
            hipLaunchKernelGGL(k_exp, dim3(grid1d((long long)img, 256)), dim3(256), 0, c.stream, amp + a * img,
                               amp + a * img, (long long)img);
            CTDM_CHECK_LAUNCH("k_exp");
        } else {
            retrieve_plain_one(c, sp, bk, H, W, cl.crop, phase + a * img, amp + a * img);
        }
        CTDM_HIP(hipStreamSynchronize(c.stream));
        unwrap_dct_plain_one(c, phase + a * img, unw + a * img, V, cl.unwrap_flags);
        CTDM_HIP(hipStreamSynchronize(c.stream));
    }
    if (cl.wrapped_opt || cl.logamp_opt) {
        const long long n = (long long)A * img;
        hipLaunchKernelGGL(k_wrap_log, dim3(grid1d(n, 256)), dim3(256), 0, c.stream, phase, amp, cl.wrapped_opt,
                           cl.logamp_opt, n);
        CTDM_CHECK_LAUNCH("k_wrap_log");
    }
    const float* ph_r = unw;
    const float* amp_r = amp;
    if (cl.recon_flags & CTDM_MAPS_MEDIAN) {  // BatchRecon: MedianFilter on both stacks, then checkMaps
        float* ph2 = (float*)c.buf("pl.phm", sizeof(float) * A * img);
        float* amp2 = (float*)c.buf("pl.ampm", sizeof(float) * A * img);
        median3x3(c, unw, ph2, A, V, V);
        median3x3(c, amp, amp2, A, V, V);
        check_maps(c, ph2, amp2, (long long)A * img);
        CTDM_HIP(hipStreamSynchronize(c.stream));
        ph_r = ph2;
        amp_r = amp2;
    }
    const EwaldGeom g = ewald_geom(cl.optics, V);
    const auto rows = angle_rows(g, cl.angles, A);
    float2* F = (float2*)c.buf("pl.F", sizeof(float2) * img * V);
    float* hits = (float*)c.buf("pl.hits", sizeof(float) * img * V);
    float2* Fm = (float2*)c.buf("pl.Fm", sizeof(float2) * img * V);
    float2* u_sp = (float2*)c.buf("pl.usp", sizeof(float2) * img);
    zero_volume(c, F, hits, V);
    for (int a = 0; a < A; ++a) {
        rytov_spectrum_plain_one(c, ph_r + a * img, amp_r + a * img, V, u_sp);
        scatter_plain_one(c, u_sp, rows[a], V, g, F, hits);
        CTDM_HIP(hipStreamSynchronize(c.stream));
    }
    normalize_measured(c, F, hits, Fm, V, g.dx);
    fft3d(c, F, V, +1);
    CTDM_HIP(hipStreamSynchronize(c.stream));
    positivity_plain(c, F, Fm, hits, (float2*)cl.ri, V, cl.iters, cl.optics, cl.recon_flags);
    CTDM_HIP(hipStreamSynchronize(c.stream));
}

void recon_from_maps(Ctx& c, const float* unwrapped, const float* map2, const ctdm_angle* ang, int A, int V,
                     const ctdm_optics& o, int iters, unsigned flags, float2* ri)
{
    CTDM_REQUIRE(unwrapped && map2 && ang && ri, "recon_from_maps: null pointer");
    CTDM_REQUIRE(A > 0 && iters >= 0, "recon_from_maps: bad sizes");
    CTDM_REQUIRE(V == 32 || V == 64 || V == 128 || V == 192 || V == 256, "recon_from_maps: volume side has no kernel");
    const size_t img = (size_t)V * V;
    const long long n = (long long)A * img;
    const bool median = flags & CTDM_MAPS_MEDIAN, amplitude = flags & CTDM_MAPS_AMPLITUDE;
    if (flags & CTDM_EXEC_PLAIN) {  // BatchRecon's own sequence: median, checkMaps, per angle copy/shift/FFT/scatter
        const EwaldGeom g = ewald_geom(o, V);
        const auto rows = angle_rows(g, ang, A);
        const float* amp_in = map2;
        if (!amplitude) {
            float* e = (float*)c.buf("rm.exp", sizeof(float) * n);
            hipLaunchKernelGGL(k_exp, dim3(grid1d(n, 256)), dim3(256), 0, c.stream, map2, e, n);
            CTDM_CHECK_LAUNCH("k_exp");
            amp_in = e;
        }
        const float* ph = unwrapped;
        const float* amp = amp_in;
        if (median || amplitude) {
            float* ph2 = (float*)c.buf("rm.ph", sizeof(float) * n);
            float* amp2 = (float*)c.buf("rm.amp", sizeof(float) * n);
            if (median) {
                median3x3(c, unwrapped, ph2, A, V, V);
                median3x3(c, amp_in, amp2, A, V, V);
            } else {
                CTDM_HIP(hipMemcpyAsync(ph2, unwrapped, sizeof(float) * n, hipMemcpyDeviceToDevice, c.stream));
                CTDM_HIP(hipMemcpyAsync(amp2, amp_in, sizeof(float) * n, hipMemcpyDeviceToDevice, c.stream));
            }
            check_maps(c, ph2, amp2, n);
            ph = ph2;
            amp = amp2;
        }
        float2* F = (float2*)c.buf("pl.F", sizeof(float2) * img * V);
        float* hits = (float*)c.buf("pl.hits", sizeof(float) * img * V);
        float2* Fm = (float2*)c.buf("pl.Fm", sizeof(float2) * img * V);
        float2* u_sp = (float2*)c.buf("pl.usp", sizeof(float2) * img);
        zero_volume(c, F, hits, V);
        for (int a = 0; a < A; ++a) {
            rytov_spectrum_plain_one(c, ph + a * img, amp + a * img, V, u_sp);
            scatter_plain_one(c, u_sp, rows[a], V, g, F, hits);
        }
        normalize_measured(c, F, hits, Fm, V, g.dx);
        fft3d(c, F, V, +1);
        positivity_plain(c, F, Fm, hits, ri, V, iters, o, flags);
        return;
    }
    const float* un = unwrapped;
    const float* la = map2;
    if (median || amplitude) {  // median, checkMaps and ln A in one pass
        float* un2 = (float*)c.buf("rm.ph", sizeof(float) * n);
        float* la2 = (float*)c.buf("rm.la", sizeof(float) * n);
        prepare_maps(c, unwrapped, map2, un2, la2, A, V, flags);
        un = un2;
        la = la2;
    }
    float2* spec = (float2*)c.buf("rm.spec", sizeof(float2) * A * img);
    rytov_spectrum(c, un, la, A, V, spec);
    double* rows = (double*)c.buf("rm.rows", sizeof(double) * 3 * (size_t)A);
    write_rows(c, rows, angle_rows(ewald_geom(o, V), ang, A));
    volume_batched(c, spec, rows, A, V, o, iters, flags, ri);
}

// Enqueues cell i of a batched call. Its maps and spectra live in the current lane's scratch, reused
// cell after cell within the lane (a lane runs its cells in order); only the crop slot and the angle
// rows, which the host writes ahead of the cells' work, are per cell.
static void enqueue_batched(Ctx& c, ctdm_cell& cl, int i, const double* rows)
{
    const int A = cl.A, V = cl.V;
    const size_t n = (size_t)A * V * V;
    float* wr = cl.wrapped_opt ? cl.wrapped_opt : (float*)c.buf("bt.wr", sizeof(float) * n);
    float* la = cl.logamp_opt ? cl.logamp_opt : (float*)c.buf("bt.la", sizeof(float) * n);
    float2* spec = (float2*)c.buf("bt.spec", sizeof(float2) * n);
    // FastUnwrapping's conventions rescale each map by its own extremes, which the reference takes of
    // calcWrapPhase's value in (-2 pi, 2 pi) (HilbertTransform.cu:3630, FastUnwrapping.cu:7): the retrieval
    // then also writes that value, and the unwrapping reads it instead of the wrapped map
    float* ph = (cl.unwrap_flags & CTDM_UNWRAP_REFCOMPAT_FAST) ? (float*)c.buf("bt.ph", sizeof(float) * n) : nullptr;
    if (cl.recon_flags & CTDM_RETRIEVE_ROWS)
        retrieve_rows(c, holo_of(cl), bg_of(cl), A, cl.H, cl.W, cl.crop.col, wr, la, ph);
    else
        retrieve_batched_dev(c, holo_of(cl), bg_of(cl), A, cl.H, cl.W, cl.V, crop_slot(c, i), wr, la, nullptr, ph);
    if (!(cl.recon_flags & CTDM_MAPS_MEDIAN) &&
        !(cl.unwrap_flags & (CTDM_UNWRAP_REFCOMPAT_FAST | CTDM_UNWRAP_REFCOMPAT_IDCT))) {
        // the unwrapping's last pass also runs the Rytov row transform; the map goes out only when asked for
        float2* P = (float2*)c.buf("ry.P", sizeof(float2) * n);
        unwrap_dct_rytov_rows(c, wr, cl.unwrapped_opt, la, A, V, cl.unwrap_flags, P);
        rytov_spectrum_cols(c, P, A, V, spec);
    } else {
        float* un = cl.unwrapped_opt ? cl.unwrapped_opt : (float*)c.buf("bt.un", sizeof(float) * n);
        unwrap_dct_batched(c, ph ? ph : wr, un, A, V, cl.unwrap_flags);
        const float* un_r = un;
        const float* la_r = la;
        if (cl.recon_flags & CTDM_MAPS_MEDIAN) {  // BatchRecon's median + checkMaps on the maps it receives
            float* un2 = (float*)c.buf("bt.unm", sizeof(float) * n);
            float* la2 = (float*)c.buf("bt.lam", sizeof(float) * n);
            prepare_maps(c, un, la, un2, la2, A, V, CTDM_MAPS_MEDIAN);
            un_r = un2;
            la_r = la2;
        }
        rytov_spectrum(c, un_r, la_r, A, V, spec);
    }
    volume_batched(c, spec, rows, A, V, cl.optics, cl.iters, cl.recon_flags, (float2*)cl.ri);
}

// Enqueues the batched sequence. With two lanes, cell i goes to lane i % 2: lane 0 is the context
// stream, lane 1 the auxiliary stream, forked from the context stream before the first cell and joined
// back after the last, so the call's work still ends on the context stream (also inside a graph
// capture, where the fork and join become graph edges). A cell's latency-bound retrieval, unwrapping
// and gridding then overlap the bandwidth-bound positivity loop of the cell on the other lane.
static void enqueue_cells(Ctx& c, ctdm_cell* cells, int ncells, int lanes, const std::vector<double*>& rows)
{
    if (lanes == 1 || ncells < 2) {
        for (int i = 0; i < ncells; ++i) enqueue_batched(c, cells[i], i, rows[i]);
        return;
    }
    c.ensure_aux();
    const hipStream_t main = c.stream;
    CTDM_HIP(hipEventRecord(c.lane_ev[0], main));
    CTDM_HIP(hipStreamWaitEvent(c.aux, c.lane_ev[0], 0));
    auto join = [&] {
        c.lane = 0;
        c.stream = main;
        const hipError_t e1 = hipEventRecord(c.lane_ev[1], c.aux);
        const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(main, c.lane_ev[1], 0) : e1;
        return e2;
    };
    try {
        for (int i = 0; i < ncells; ++i) {
            c.lane = i & 1;
            c.stream = c.lane ? c.aux : main;
            enqueue_batched(c, cells[i], i, rows[i]);
        }
    } catch (...) {
        join();  // a capture can only end once lane 1 has joined back
        throw;
    }
    CTDM_HIP(join());
}

// The captured sequence depends on pointers, shapes, flags and optics (launch arguments). Crops (except
// the row-wise retrieval's column, a launch argument) and angle rows are read from device tables that
// are rewritten before every launch, so their values are not part of the key.
static GraphKey graph_key(const ctdm_cell* cells, int ncells)
{
    GraphKey k;
    for (int i = 0; i < ncells; ++i) {
        const ctdm_cell& cl = cells[i];
        const bool rows = cl.recon_flags & CTDM_RETRIEVE_ROWS;
        for (uintptr_t w : {(uintptr_t)cl.holo, (uintptr_t)cl.bg, (uintptr_t)cl.ri, (uintptr_t)cl.wrapped_opt,
                            (uintptr_t)cl.unwrapped_opt, (uintptr_t)cl.logamp_opt, (uintptr_t)cl.A,
                            (uintptr_t)cl.H, (uintptr_t)cl.W, (uintptr_t)cl.V, (uintptr_t)cl.iters,
                            (uintptr_t)cl.unwrap_flags, (uintptr_t)cl.recon_flags,
                            (uintptr_t)(rows ? cl.crop.row : 0), (uintptr_t)(rows ? cl.crop.col : 0),
                            (uintptr_t)(rows ? cl.crop.radius : 0)})
            k.words.push_back(w);
        uint64_t bits[3];
        std::memcpy(&bits[0], &cl.optics.wavelength, 8);
        std::memcpy(&bits[1], &cl.optics.n_med, 8);
        std::memcpy(&bits[2], &cl.optics.dx, 8);
        for (uint64_t b : bits) k.words.push_back((uintptr_t)b);
    }
    return k;
}

void reconstruct(Ctx& c, ctdm_cell* cells, int ncells, const ctdm_batch_opts& opts)
{
    CTDM_REQUIRE(ncells >= 0, "reconstruct: negative cell count");
    for (int i = 0; i < ncells; ++i) validate_cell(cells[i]);
    if (opts.mode == 0) {
        for (int i = 0; i < ncells; ++i) run_plain(c, cells[i]);
        return;
    }
    CTDM_REQUIRE(opts.mode == 1, "reconstruct: mode must be 0 (plain) or 1 (batched)");
    CTDM_REQUIRE(opts.lanes >= 0 && opts.lanes <= 2, "reconstruct: lanes must be 0, 1 or 2");
    // Crops, all before any cell's work and never captured. A given crop is validated and written to
    // the cell's slot; a row-wise cell locates its column synchronously (frame 1, HilbertTransform.cu:1634).
    // The others queue the sideband search (frame 1, HilbertTransform.cu:1572), which writes the slot on
    // the device: the host reads the keys back only after the cells' work is queued (finish_crops), so
    // the GPU goes from the searches straight into the cells without waiting for the host.
    std::vector<int> located;
    for (int i = 0; i < ncells; ++i)
        if (!(cells[i].recon_flags & CTDM_RETRIEVE_ROWS) && cells[i].crop.radius < 0) located.push_back(i);
    unsigned long long* keys = nullptr;
    unsigned long long* keys_dev = nullptr;
    if (!located.empty()) {
        keys = c.host_words((int)located.size());
        CTDM_HIP(hipHostGetDevicePointer((void**)&keys_dev, keys, 0));
    }
    for (int i = 0, k = 0; i < ncells; ++i) {
        ctdm_cell& cl = cells[i];
        if (cl.recon_flags & CTDM_RETRIEVE_ROWS) {
            cl.crop = cell_crop(c, cl);
        } else if (cl.crop.radius >= 0) {
            validate_crop(cl.H, cl.W, cl.crop);
            set_crop(c, crop_slot(c, i), cl.crop, cl.H, cl.W);
        } else {
            enqueue_sideband(c, bg_of(cl), cl.H, cl.W, cl.V, 0, crop_slot(c, i), keys_dev + k++);
        }
    }
    if (!located.empty()) {
        if (!c.wait_ev) CTDM_HIP(hipEventCreateWithFlags(&c.wait_ev, hipEventDisableTiming));
        CTDM_HIP(hipEventRecord(c.wait_ev, c.stream));
    }
    // Reports the located crops (waiting for the searches only) and rejects one whose window leaves the
    // spectrum; the cells' work is queued by then and reads the same crop, always in bounds.
    auto finish_crops = [&] {
        if (located.empty()) return;
        c.spin_event(c.wait_ev);
        for (size_t k = 0; k < located.size(); ++k) {
            ctdm_cell& cl = cells[located[k]];
            cl.crop = decode_sideband(keys[k], cl.H, cl.W, cl.V);
        }
        for (int i : located) validate_crop(cells[i].H, cells[i].W, cells[i].crop);
    };
    if (opts.autotune) {
        std::vector<int> sizes;
        for (int i = 0; i < ncells; ++i) sizes.push_back(cells[i].V);
        autotune(c, sizes.data(), (int)sizes.size(), false);
    }
    // Angle rows of every cell, written ahead of the cells' work on the context stream (before lane 1 is
    // forked from it), into the tables the launch sequence reads.
    std::vector<std::vector<AngleRow>> host_rows(ncells);
    for (int i = 0; i < ncells; ++i)
        host_rows[i] = angle_rows(ewald_geom(cells[i].optics, cells[i].V), cells[i].angles, cells[i].A);
    GraphKey key;
    std::map<GraphKey, GraphEntry>::iterator it = c.graphs.end();
    if (opts.graph) {
        key = graph_key(cells, ncells);
        key.words.push_back((uintptr_t)(opts.lanes == 1 ? 1 : 2));
        key.words.push_back((uintptr_t)c.tune_epoch);  // a graph holds the kernel choices made when captured
        it = c.graphs.find(key);
    }
    if (it != c.graphs.end()) {  // replay: the graph's own row tables get this call's angles
        GraphEntry& ge = it->second;
        for (int i = 0; i < ncells; ++i) write_rows(c, ge.rows[i], host_rows[i], true);
        ge.last_use = ++c.graph_clock;
        CTDM_HIP(hipGraphLaunch(ge.exec, c.stream));
        finish_crops();
        return;
    }
    std::vector<double*> rows(ncells);
    for (int i = 0; i < ncells; ++i) {
        rows[i] = rows_slot(c, i, cells[i].A);
        write_rows(c, rows[i], host_rows[i], true);
    }
    if (!opts.graph) {
        enqueue_cells(c, cells, ncells, opts.lanes, rows);
        finish_crops();
        return;
    }
    // Eager pass: allocates every buffer and table the sequence needs and produces the result.
    enqueue_cells(c, cells, ncells, opts.lanes, rows);
    CTDM_HIP(hipStreamSynchronize(c.stream));
    c.trim_graphs(kMaxGraphs - 1);  // room for the new graph; frees buffers only evicted graphs held
    hipStream_t cs = nullptr;
    CTDM_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipStream_t saved = c.stream;
    c.stream = cs;
    c.capturing = true;
    GraphEntry ge;
    try {
        CTDM_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_cells(c, cells, ncells, opts.lanes, rows);
        } catch (...) {
            hipGraph_t g = nullptr;
            hipStreamEndCapture(cs, &g);
            if (g) hipGraphDestroy(g);
            throw;
        }
        CTDM_HIP(hipStreamEndCapture(cs, &ge.graph));
        CTDM_HIP(hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0));
    } catch (...) {
        c.capturing = false;
        c.stream = saved;
        hipStreamDestroy(cs);
        throw;
    }
    c.capturing = false;
    c.stream = saved;
    CTDM_HIP(hipStreamDestroy(cs));
    // The graph holds every scratch pointer it was captured with (a superset of what it reads): such a
    // buffer is retired rather than freed if it grows while the graph is cached.
    for (const auto& kv : c.scratch)
        if (kv.second.ptr) ge.ptrs.push_back(kv.second.ptr);
    std::sort(ge.ptrs.begin(), ge.ptrs.end());
    ge.rows = rows;
    ge.last_use = ++c.graph_clock;
    c.graphs[key] = ge;
    finish_crops();
}

}  // namespace ctdm
