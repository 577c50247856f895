// Host-side internals shared by the libctdm translation units: the context object, status
// plumbing, twiddle tables and the kernel-launch entry points each module exports to the API layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ctdm.h"

namespace ctdm {

// Thrown inside the library, converted to a ctdm_status at the C boundary (never crosses it).
struct Error {
    ctdm_status status;
    std::string msg;
};

[[noreturn]] inline void fail(ctdm_status s, const std::string& msg) { throw Error{s, msg}; }

#define CTDM_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            ::ctdm::fail(e_ == hipErrorOutOfMemory ? CTDM_ENOMEM : CTDM_EDEVICE,                \
                         std::string(#call) + ": " + hipGetErrorString(e_));                    \
    } while (0)

#define CTDM_CHECK_LAUNCH(name)                                                                 \
    do {                                                                                        \
        hipError_t e_ = hipGetLastError();                                                      \
        if (e_ != hipSuccess)                                                                   \
            ::ctdm::fail(CTDM_EDEVICE, std::string("launch ") + (name) + ": " +                 \
                                           hipGetErrorString(e_));                              \
    } while (0)

#define CTDM_REQUIRE(cond, msg)                                                                 \
    do {                                                                                        \
        if (!(cond)) ::ctdm::fail(CTDM_EINVAL, msg);                                            \
    } while (0)

// Device buffer owned by the context; grows on demand (never inside graph capture). A buffer that a
// cached graph references is retired instead of freed when it grows, and freed once the last graph
// referencing it has been evicted.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct Ctx;

// Per-size kernel choices picked by the autotuner (defaults are the untuned picks).
struct Tune {
    int pencils[4] = {16, 16, 16, 8};  // pencils per workgroup of the x, y, z+dc and gridding passes
    int rev[2] = {0, 0};               // x and y passes sweep the z planes backwards
    int yplan = 0;                     // radix plan of the y pass: 0 Plan<V>, 1 PlanAlt<V>
    bool tuned = false;
    double ms[4][2] = {};               // measured ms per launch for B = 8, 16 (after tuning)
    double sweep_ms[3] = {};            // loop iteration (x + 2 y + z+dc) for sweeps (fwd, fwd), (rev, fwd), (fwd, rev)
    double yplan_ms[2] = {};            // y pass per launch with either radix plan
    double grid_ms[2] = {};             // gridding pass per launch with 8 / 16 pencils
};

struct GraphKey {
    std::vector<uintptr_t> words;
    bool operator<(const GraphKey& o) const { return words < o.words; }
};

struct GraphEntry {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void*> ptrs;     // sorted scratch pointers the captured sequence references
    std::vector<double*> rows;   // per cell: the angle-row table the graph reads (rewritten before each replay)
    uint64_t last_use = 0;       // LRU stamp
};

// Captured graphs kept per context; the least recently used one is evicted beyond this.
constexpr size_t kMaxGraphs = 8;

struct Ctx {
    hipEvent_t tev[2] = {nullptr, nullptr};  // when set: recorded at start and end of the next volume pass
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    std::map<int, float2*> twiddles;   // N -> device table exp(-2*pi*i*m/N), m < N
    std::map<std::string, DevBuf> scratch;
    std::map<int, Tune> tune;
    uint64_t tune_epoch = 0;  // bumped by every tuning change; part of the graph key (graphs bake in choices)
    std::map<GraphKey, GraphEntry> graphs;
    uint64_t graph_clock = 0;
    std::vector<void*> retired;  // grown buffers still referenced by a cached graph
    // Host copy of what each per-cell angle-row table holds, so that a replay with unchanged angles issues
    // no copy. Entries are dropped when their buffer is freed.
    std::map<const void*, std::vector<double>> table_contents;
    void release(void* p);  // frees p (after the caller synchronised) and forgets its table contents
    // True when a cached graph references device pointer p.
    bool referenced(const void* p) const;
    // Evicts least recently used graphs until at most `keep` remain, then frees retired buffers that no
    // remaining graph references (synchronises the context stream first).
    void trim_graphs(size_t keep);
    bool capturing = false;
    float* noalloc_scratch = nullptr;  // this device's address of the allocation-free entry point's array
    int max_lds = 65536;
    int cus = 0;  // compute units of the device (0: unknown)
    int dc_V = 0;    // volume side the cached data-consistency list ("dc.*") was built for
    int dc_cap = 0;  // its entries per pencil

    unsigned long long* host_words_ = nullptr;  // small pinned host buffer (hipHostMalloc)

    // Lanes of a batched multi-cell sequence: cells alternate between lane 0 (the context stream) and
    // lane 1 (an auxiliary stream forked from it and joined back). Scratch named while lane 1 is
    // active is lane 1's own copy, so the two lanes never share a buffer.
    int lane = 0;
    hipStream_t aux = nullptr;                    // lane 1 stream (default priority), made on first use
    hipEvent_t lane_ev[2] = {nullptr, nullptr};   // fork and join events of lane 1
    void ensure_aux();

    const float2* tw(int n);
    // Scratch buffer `name` of at least `bytes` bytes (contents undefined). Lane 1 has its own copy
    // unless `shared` (buffers written only on the context stream outside the lanes' work: read-only
    // tables, per-cell crop slots).
    void* buf(const std::string& name, size_t bytes, bool shared = false);
    // Pinned host words for small results read back from the device (contents undefined; a larger
    // request replaces the buffer, so no device write to it may be pending).
    unsigned long long* host_words(int n);
    int host_words_n_ = 0;
    // Waits for the context stream by polling it: a host-synchronous search returns as soon as its last
    // command has finished, without the wake-up latency of a blocking synchronisation.
    void spin_sync();
    // The same for an event (made on first use): waits for the work recorded before it only.
    hipEvent_t wait_ev = nullptr;
    void spin_event(hipEvent_t e);
    // Tuning entry of side n; a side this context has not tuned starts from the choices the autotuner
    // made for it on a device of the same architecture earlier in the process, if any.
    Tune& tuning(int n);
    std::string arch;  // gcnArchName of the context's device
};

// ---------------------------------------------------------------- module launchers
// Retrieval (k_retrieve.hip)
// Hologram pixels as the caller stores them: 32-bit floats, or 8- or 16-bit unsigned integers (camera
// frames; the reference reads 8-bit BMP frames and widens them on the host, HilbertTransform.cu:419-420,
// 660-661, 2694). The kernels that read holograms widen integer pixels exactly as they load them, so an
// integer frame gives the same results as the same values passed as floats.
enum class Px : int { F32 = 0, U8 = 1, U16 = 2 };
struct Pixels {
    const void* p = nullptr;
    Px px = Px::F32;
    size_t size() const { return px == Px::F32 ? 4 : (px == Px::U16 ? 2 : 1); }
    // The pixels n pixels further on (image a of a stack: offset(a * H * W)).
    Pixels offset(size_t n) const { return Pixels{p ? (const char*)p + n * size() : nullptr, px}; }
    explicit operator bool() const { return p != nullptr; }
};
inline Pixels f32_pixels(const float* p) { return Pixels{p, Px::F32}; }
// Crop of the batched retrieval as its kernels read it from device memory, so that one launch sequence
// (and one captured graph) serves any carrier position: the crop itself and the passband it implies
// (band [lo, lo + n) of offsets, the same for rows and columns; k_lo = col - W/2 + lo; r0 = row - H/2).
struct CropDev {
    int row, col, radius, M;
    int lo, n, k_lo, r0;
};
// Writes `slot` from a host crop with a one-thread kernel (no copy node, so it may be captured).
void set_crop(Ctx& c, CropDev* slot, const ctdm_crop& cr, int H, int W);
// Queues the sideband search of one background image without waiting: the crop goes to `slot` (if not
// null) and the 64-bit search key to `key_dev` (device-visible pinned host memory).
void enqueue_sideband(Ctx& c, Pixels bg, int H, int W, int M, int row_start, CropDev* slot,
                      unsigned long long* key_dev);
// The crop a search key stands for (obtainRadius's row, column and radius).
ctdm_crop decode_sideband(unsigned long long key, int H, int W, int M);
void validate_crop(int H, int W, const ctdm_crop& cr);
void find_sideband(Ctx& c, Pixels bg, int H, int W, int M, int row_start, ctdm_crop* out);
int find_sideband_rows(Ctx& c, Pixels bg, int H, int W);
void angle_from_image(Ctx& c, Pixels img, int batch, int H, int W, double pixel_um, double mag,
                      double wavelength_nm, double n_med, double* angles);
// refph (may be null): calcWrapPhase's own value atan2(BG) - atan2(SP) in (-2 pi, 2 pi), the input of
// FastUnwrapping's min-max rescaling in the reference (HilbertTransform.cu:3630; FastUnwrapping.cu:7).
void retrieve_rows(Ctx& c, Pixels holo, Pixels bg, int batch, int H, int W, int center, float* wrapped,
                   float* logamp, float* refph = nullptr);
void retrieve_batched(Ctx& c, Pixels holo, Pixels bg, int batch, int H, int W, const ctdm_crop& crop,
                      float* wrapped, float* logamp, float2* field);
// The same with the crop read from device memory (crop size M fixed on the host).
void retrieve_batched_dev(Ctx& c, Pixels holo, Pixels bg, int batch, int H, int W, int M, const CropDev* cd,
                          float* wrapped, float* logamp, float2* field, float* refph = nullptr);
// Reference-structured retrieval of one hologram pair (full 2-D FFTs, shift, crop, inverse,
// scale+conjugate, calcWrapPhase semantics: phase in (-2pi, 2pi), amplitude ratio).
void retrieve_plain_one(Ctx& c, Pixels holo, Pixels bg, int H, int W, const ctdm_crop& crop, float* phase,
                        float* amp);

// Generic FFT passes (k_fft.hip)
// Batched 1-D transforms of length n over pencils of a strided array (see k_fft.hip).
void fft_rows(Ctx& c, float2* data, int n, long long count, int dir, float scale, bool conj_out);
void fft_cols(Ctx& c, float2* data, int n, int width, long long stride_n, long long nplanes,
              long long stride_plane, int dir, float scale, bool conj_out);
void fftshift2d(Ctx& c, float2* data, int W, int H, long long batch);
void fftshift3d(Ctx& c, float2* data, int V);

// Unwrapping (k_unwrap.hip)
void unwrap_dct_batched(Ctx& c, const float* wrapped, float* unwrapped, int batch, int M,
                        unsigned flags);
void unwrap_dct_plain_one(Ctx& c, const float* wrapped, float* unwrapped, int M, unsigned flags);
void unwrap_dct_adjoint(Ctx& c, const float* wrapped, const float* g_out, float* g_in, int batch,
                        int M, unsigned flags);
// unwrap_dct_batched followed by the row pass of rytov_spectrum in one sequence: the last unwrapping
// pass also transforms each row of u = logamp + i phi into P (batch M x M complex); `unwrapped` may be
// null. Flags without CTDM_UNWRAP_REFCOMPAT_FAST and CTDM_UNWRAP_REFCOMPAT_IDCT.
void unwrap_dct_rytov_rows(Ctx& c, const float* wrapped, float* unwrapped, const float* logamp, int batch, int M,
                           unsigned flags, float2* P);
// Allocation-free entry point: device scratch address per device and the unwrap tables (context set-up),
// then solves with the fixed scratch array (no allocation, no table building).
void noalloc_setup(Ctx& c);
void unwrap_dct_noalloc(Ctx& c, const float* wrapped, float* unwrapped, int batch, int M, unsigned flags);

// Goldstein (k_goldstein.hip)
void goldstein(Ctx& c, const float* wrapped, int batch, int H, int W, int8_t* res, uint8_t* cuts,
               int32_t* k2pi, float* unwrapped);

// Reconstruction (k_recon.hip)
struct EwaldGeom {
    double fm0, fm02, df, dx, k2, n_med;
};
EwaldGeom ewald_geom(const ctdm_optics& o, int V);
struct AngleRow {
    double fx0, fy0, fz0;
};
std::vector<AngleRow> angle_rows(const EwaldGeom& g, const ctdm_angle* ang, int A);

void rytov_spectrum(Ctx& c, const float* unwrapped, const float* logamp, int A, int M,
                    float2* spec);
// The column pass of rytov_spectrum alone, from the row spectra P that unwrap_dct_rytov_rows wrote.
void rytov_spectrum_cols(Ctx& c, const float2* P, int A, int M, float2* spec);
void rytov_spectrum_plain_one(Ctx& c, const float* phase, const float* amp, int M, float2* u_sp);
void ewald_scatter(Ctx& c, const float2* spec, const ctdm_angle* ang, int A, int M, int V,
                   const ctdm_optics& o, float2* F, float* hits);
void ewald_gather(Ctx& c, const float2* F, const ctdm_angle* ang, int A, int M, int V,
                  const ctdm_optics& o, float2* spec);
void normalize_measured(Ctx& c, float2* F, const float* hits, float2* Fm, int V, double dx);
void fft3d(Ctx& c, float2* vol, int V, int dir);
void positivity_plain(Ctx& c, float2* f, const float2* Fm, const float* hits, float2* n, int V,
                      int iters, const ctdm_optics& o, unsigned flags);
void positivity_fused(Ctx& c, float2* f, const float2* Fm, const float* hits, float2* n, int V,
                      int iters, const ctdm_optics& o, unsigned flags);
// Batched cell volume: spectra of all angles -> fused gridding + inverse FFT + positivity loop.
// `rows` is the cell's device table of A angle rows (fx0, fy0, fz0), written by write_rows.
void volume_batched(Ctx& c, const float2* spec, const double* rows, int A, int V,
                    const ctdm_optics& o, int iters, unsigned flags, float2* n_out);
// Copies host angle rows to the device table `dst` (3 doubles per angle) on the context stream; never
// under capture (a host copy must not be recorded). With `cached`, the copy is skipped when `dst`
// already holds these rows (only for tables written through this function alone).
void write_rows(Ctx& c, double* dst, const std::vector<AngleRow>& rows, bool cached = false);
// objective: HOST array of iters values (may be null): J before each step; synchronises when given.
void fit_pgd(Ctx& c, float2* f, const float2* spec, const ctdm_angle* ang, int A, int V,
             const ctdm_optics& o, int iters, double step, float2* n, unsigned flags, double* objective = nullptr);

// Average ms of one launch of volume pass `kind` (0 x, 1 y, 2 z+dc) with B pencils per workgroup.
double time_pass(Ctx& c, int kind, int V, int B, int reps, double* bytes);
// Average ms of one gridding pass at side V with B pencils per workgroup for A synthetic angles.
double time_grid(Ctx& c, int V, int B, int A, int reps);
void time_loop(Ctx& c, int V, int iters, double ms[3], double bytes[3]);

// Pipeline (pipeline.cpp)
void reconstruct(Ctx& c, ctdm_cell* cells, int ncells, const ctdm_batch_opts& opts);
void recon_from_maps(Ctx& c, const float* unwrapped, const float* logamp, const ctdm_angle* ang, int A, int V,
                     const ctdm_optics& o, int iters, unsigned flags, float2* ri);
void median3x3(Ctx& c, const float* in, float* out, int batch, int H, int W);
void check_maps(Ctx& c, float* phase, float* amp, long long n);
void prepare_maps(Ctx& c, const float* ph, const float* m2, float* ph_out, float* la_out, int batch, int M,
                  unsigned flags);
std::string autotune(Ctx& c, const int* sizes, int nsizes, bool force);

inline int grid1d(long long n, int block) { return (int)((n + block - 1) / block); }

}  // namespace ctdm
