/*
 * ctdm.h — C ABI of the MI355X-native cTDM reconstruction library (libctdm.so).
 *
 * Path: interferogram stack + background holograms -> wrapped / unwrapped phase + log-amplitude
 *       -> Rytov spectra -> Ewald-cap k-space -> positivity-constrained 3-D RI volume.
 *
 * Conventions shared by every entry point
 *   - All array pointers are caller-owned DEVICE memory unless a comment says "host".
 *   - Images are row-major with idx = x + W*y; volumes idx = x + V*y + V*V*z (the reference's
 *     layout, e.g. HilbertTransform.cu:3566-3586, Recon_kernel.cu:262).
 *   - Complex values are interleaved float pairs (ctdm_complex), bit-compatible with cufftComplex.
 *   - Every call is stream-ordered on the context's stream and returns a ctdm_status; nothing
 *     calls exit() (the reference's checkCudaErrors does, HilbertTransform.cu:480-574).
 *     ctdm_find_sideband and ctdm_autotune additionally synchronise the stream because their
 *     results are host values.
 *   - The context owns scratch memory, FFT twiddle tables, the autotune cache and captured graphs.
 *     One context must not be used by two host threads at the same time.
 */
#ifndef CTDM_H
#define CTDM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ctdm_status {
    CTDM_OK = 0,
    CTDM_EINVAL = 1,       /* bad argument (sizes, null pointers, crop outside the spectrum) */
    CTDM_ENOMEM = 2,       /* device allocation failed */
    CTDM_EDEVICE = 3,      /* HIP runtime / kernel launch error */
    CTDM_EUNSUPPORTED = 4  /* size or flag combination without a kernel */
} ctdm_status;

typedef struct ctdm_ctx ctdm_ctx;

typedef struct ctdm_complex { float x, y; } ctdm_complex;

/* hipStream_t, passed opaquely so that FFI callers need no HIP headers; NULL = default stream. */
typedef void *ctdm_stream;

typedef struct ctdm_config {
    int device;        /* HIP device ordinal the context binds to */
    int reserved[7];   /* must be zero */
} ctdm_config;

/* Sideband crop. Reference: obtainRadius() (HilbertTransform.cu:3425-3455) returns (r, c, radius);
 * get1stOrder_new() (HilbertTransform.cu:3566-3586) crops an M x M window (M = H/4 there). */
typedef struct ctdm_crop {
    int row;     /* sideband centre row in the fftshifted H x W spectrum */
    int col;     /* sideband centre column in the fftshifted spectrum */
    int radius;  /* half-width of the square passband, |row - H/2| / 3 in the reference */
    int M;       /* crop side; also the Rytov spectrum side and the volume side */
} ctdm_crop;

/* Optics of one cell. Reference constants: BatchRecon (HilbertTransform.cu:1840-1862). */
typedef struct ctdm_optics {
    double wavelength;  /* vacuum wavelength [m] */
    double n_med;       /* medium refractive index */
    double dx;          /* sampling of the cropped field = voxel pitch [m]
                           (reference: CameraPixelSize/Mag * 1024 / ReconZ) */
} ctdm_optics;

/* Illumination angle of one hologram [rad] (reference sampleAngleRadX/Y; EdwardSphere takes
 * f_x0 from ang_y and f_y0 from ang_x, Recon_kernel.cu:175-176 — reproduced as is). */
typedef struct ctdm_angle { double ang_x, ang_y; } ctdm_angle;

/* ---- unwrapping flags (ctdm_unwrap_dct) ---- */
#define CTDM_UNWRAP_WRAPDIFF        0x0u  /* RHS = divergence of wrapped forward differences (DCT_Unwrapping.cu:336-417) */
#define CTDM_UNWRAP_SINCOS          0x1u  /* RHS = cos(phi) L[sin(phi)] - sin(phi) L[cos(phi)] (FastUnwrapping.cu:37-96 idea) */
#define CTDM_UNWRAP_REFCOMPAT_EIG   0x2u  /* eigenvalues 2cos(4*pi*i/(4N-1)) + 2cos(4*pi*j/(4N-1)) - 4 (DCT_Unwrapping.cu:310) */
#define CTDM_UNWRAP_REFCOMPAT_IDCT  0x4u  /* (i+1) twiddle of the reference IDCT (DCT_Unwrapping.cu:230-231) */
#define CTDM_UNWRAP_CONGRUENT       0x8u  /* phi_w + 2*pi*round((phi_ls - phi_w)/2*pi) (FastUnwrapping.cu:333-344) */
#define CTDM_UNWRAP_REFCOMPAT_FAST  0x10u /* FastUnwrapping's conventions (FastUnwrapping.cu:7,108,119-142,346-369):
                                              each input map rescaled to 2*pi*(x - min)/(max - min) first
                                              (PhaseNormalize), each output map minus its minimum (subtractMin) */
#define CTDM_UNWRAP_REFCOMPAT       (CTDM_UNWRAP_REFCOMPAT_EIG | CTDM_UNWRAP_REFCOMPAT_IDCT)

/* ---- reconstruction flags ---- */
#define CTDM_DC_HITMASK     0x0u   /* data consistency where the hit count is > 0 (default) */
#define CTDM_DC_REFCOMPAT   0x1u   /* reference test: F2.re != 0 && F2.im != 0 (Recon_kernel.cu:406) */
#define CTDM_POS_NO_SHIFT   0x2u   /* skip the final 3-D octant shift (HilbertTransform.cu:1974) */
#define CTDM_MAPS_MEDIAN    0x4u   /* 3x3 median of the phase and amplitude maps before gridding
                                      (MedianFilter, Recon_kernel.cu:7-27; see ctdm_median3x3) */
#define CTDM_MAPS_AMPLITUDE 0x8u   /* ctdm_recon_from_maps: the second map holds the amplitude A (BatchRecon's
                                      AmpStack), sanitised by checkMaps and logged on the device */
#define CTDM_RETRIEVE_ROWS  0x10u  /* ctdm_reconstruct: row-wise retrieval (extractQPI_AlgorithmB, the
                                      reference's default) instead of the 2-D crop; crop.col holds the
                                      sideband column (crop.radius < 0: located from bg[0]); V = W / 4 */
#define CTDM_HOLO_U8        0x20u  /* ctdm_cell: holo and bg hold 8-bit unsigned pixels (the reference's BMP
                                      frames, HilbertTransform.cu:419-420, 660-661) instead of float32 */
#define CTDM_HOLO_U16       0x40u  /* ctdm_cell: holo and bg hold 16-bit unsigned pixels */
#define CTDM_EXEC_PLAIN     0x100u /* run the reference-structured, unfused kernel sequence */

/* Pixel format of hologram, background and angle-calibration images (the *_px entry points). Integer
 * pixels are widened to float exactly as they are loaded, so results equal those for the same values
 * passed as float32; the reference reads 8-bit BMP frames (read_bmp, HilbertTransform.cu:2694) and
 * widens them on the host (:660-661). The 2-D retrieval reads pixels in pairs: its image pointers must
 * be aligned to two pixels (CTDM_EINVAL otherwise). */
typedef enum ctdm_pixel {
    CTDM_PIXEL_F32 = 0,
    CTDM_PIXEL_U8 = 1,
    CTDM_PIXEL_U16 = 2
} ctdm_pixel;

/* ---------------------------------------------------------------- context */
ctdm_status ctdm_create(ctdm_ctx **ctx, const ctdm_config *cfg, ctdm_stream stream);
ctdm_status ctdm_destroy(ctdm_ctx *ctx);
ctdm_status ctdm_set_stream(ctdm_ctx *ctx, ctdm_stream stream);
ctdm_status ctdm_sync(ctdm_ctx *ctx);
/* Text of the last error raised on this context (never NULL). */
const char *ctdm_last_error(const ctdm_ctx *ctx);
/* Library version string (host). */
const char *ctdm_version(void);

/* ---------------------------------------------------------------- retrieval (L2) */
/* Replaces obtainRadius (HilbertTransform.cu:3425-3455) applied to the fftshifted spectrum of the
 * background (HilbertTransform.cu:1564-1572): argmax of |BG spectrum| over shifted rows
 * [row_start, H), raster order, first maximum wins; radius = |row - H/2| / 3. row_start <= 0 selects
 * the reference's 600*H/1024. Writes the crop to HOST memory *out (synchronises the stream). */
ctdm_status ctdm_find_sideband(ctdm_ctx *ctx, const float *bg, int H, int W, int M, int row_start,
                               ctdm_crop *out);

/* Replaces extractQPI_AlgorithmA (HilbertTransform.cu:1546-1585) + calcWrapPhase
 * (HilbertTransform.cu:3618-3639) for `batch` hologram pairs (holo, bg: batch x H x W float32).
 * Outputs per item, M x M: wrapped = arg(SP) - arg(BG) wrapped to (-pi, pi] (the reference value
 * mod 2*pi), logamp = ln(|SP| / |BG|). field_opt (may be NULL): SP/BG complex ratio.
 * Non-finite results are replaced by 0 as in calcWrapPhase. */
ctdm_status ctdm_retrieve(ctdm_ctx *ctx, const float *holo, const float *bg, int batch, int H, int W,
                          const ctdm_crop *crop, float *wrapped, float *logamp,
                          ctdm_complex *field_opt);

/* Replaces the sideband search of extractQPI_AlgorithmB (sequence1DFFT frame 1,
 * HilbertTransform.cu:1634-1643): every row of the background, resampled to H/4 rows as
 * bilinear_interpolation_kernel does (:3658-3718), is transformed along x; HistogramFT (:1728-1748) sums
 * |X| over the rows per fftshifted column; the first maximum over columns [0.6 W, W) is the sideband
 * column, mirrored to W - column when it lies within W/8 of the end. Writes the column to HOST *center
 * (synchronises). Square 256, 512 or 1024 holograms. */
ctdm_status ctdm_find_sideband_rows(ctdm_ctx *ctx, const float *bg, int H, int W, int *center);

/* Replaces extractQPI_AlgorithmB (HilbertTransform.cu:1587-1605) + sequence1DFFT (:1608-1665) +
 * calcWrapPhase (:3618-3639), the reference's default retrieval (QPI_Method 0), for `batch` hologram
 * pairs: bilinear resampling to H/4 rows, forward transform of each row, the W/4 fftshifted columns
 * [center - W/8, center + W/8) kept and swapped by halves (CropFTdomain, shiftArray), inverse W/4-point
 * transform. Outputs per item, (W/4) x (H/4) maps in the reference's transposed layout (copySharedMem
 * :1704-1726): value (n, y') at n * (H/4) + y', with y' the resampled row and n the filtered column;
 * wrapped = arg(SP) - arg(BG) in (-pi, pi], logamp = ln(|SP| / |BG|), non-finite -> 0. */
ctdm_status ctdm_retrieve_rows(ctdm_ctx *ctx, const float *holo, const float *bg, int batch, int H, int W,
                               int center, float *wrapped, float *logamp);

/* Replaces AngCal_GPU (HilbertTransform.cu:2920-2982), the illumination angle of an angle-calibration
 * image: 2-D DFT of each of `batch` H x W images, fftshift, log10|F| (non-finite -> 0), first maximum over
 * the flat shifted indices [0.53 HW, 0.79 HW); with posX its column,
 * fx = (W/2 - 1 - posX) / (W * pixel_um * 1e-6), sin(theta) = mag * fx / (n_med / (wavelength_nm * 1e-9)),
 * angle = |asin| with the opposite sign of sin(theta) (the reference's convention; W = 1024 gives its
 * literal 511 and 1024). Writes `batch` angles [rad] to HOST memory (synchronises). Square 256, 512 or 1024
 * images. */
ctdm_status ctdm_angle_from_image(ctdm_ctx *ctx, const float *img, int batch, int H, int W, double pixel_um,
                                  double mag, double wavelength_nm, double n_med, double *angles);

/* The same five entry points for images of any ctdm_pixel format (holo, bg and img point to pixels of
 * that format; the float32 entry points above are these with CTDM_PIXEL_F32). */
ctdm_status ctdm_find_sideband_px(ctdm_ctx *ctx, const void *bg, ctdm_pixel px, int H, int W, int M,
                                  int row_start, ctdm_crop *out);
ctdm_status ctdm_retrieve_px(ctdm_ctx *ctx, const void *holo, const void *bg, ctdm_pixel px, int batch, int H,
                             int W, const ctdm_crop *crop, float *wrapped, float *logamp, ctdm_complex *field_opt);
ctdm_status ctdm_find_sideband_rows_px(ctdm_ctx *ctx, const void *bg, ctdm_pixel px, int H, int W, int *center);
ctdm_status ctdm_retrieve_rows_px(ctdm_ctx *ctx, const void *holo, const void *bg, ctdm_pixel px, int batch,
                                  int H, int W, int center, float *wrapped, float *logamp);
ctdm_status ctdm_angle_from_image_px(ctdm_ctx *ctx, const void *img, ctdm_pixel px, int batch, int H, int W,
                                     double pixel_um, double mag, double wavelength_nm, double n_med,
                                     double *angles);

/* ---------------------------------------------------------------- unwrapping (L3) */
/* Replaces DCT_UWLS_Unwrapped (DCT_Unwrapping.cu:10-59; RHS option from FastUnwrapping.cu:5-117):
 * least-squares unwrapping with Neumann boundaries via an orthonormal DCT-II Poisson solve, for
 * `batch` M x M maps. flags: CTDM_UNWRAP_*. */
ctdm_status ctdm_unwrap_dct(ctdm_ctx *ctx, const float *wrapped, float *unwrapped, int batch, int M,
                            unsigned flags);

/* Allocation-free variant of ctdm_unwrap_dct. It has no counterpart in the reference: it is for callers
 * that cannot allocate device memory during the call (code under hipGraph stream capture, tight serving
 * loops, harnesses that forbid hipMalloc). Its two intermediate spectra live in a fixed device array of
 * 2 x CTDM_NOALLOC_MAX_ELEMS floats (32 MiB) that every GPU the library runs on holds until the process
 * ends; ctdm_create sets up that array's address and builds the tables the call reads, so the call itself
 * allocates nothing. batch * M * M above CTDM_NOALLOC_MAX_ELEMS is rejected with CTDM_EINVAL. All calls
 * on one GPU share the array (it keeps the last call's contents, which every call overwrites before
 * reading): calls that overlap on one GPU, from two streams or two contexts, race. Same flags, kernels
 * and results as ctdm_unwrap_dct (CTDM_EXEC_PLAIN and CTDM_UNWRAP_REFCOMPAT_FAST, which needs a
 * third buffer, are not accepted). */
#define CTDM_NOALLOC_MAX_ELEMS (1u << 22)
ctdm_status ctdm_unwrap_dct_noalloc(ctdm_ctx *ctx, const float *wrapped, float *unwrapped, int batch, int M,
                                    unsigned flags);

/* Adjoint (transpose) of the Jacobian of ctdm_unwrap_dct at `wrapped` (no reference counterpart):
 * grad_wrapped = J^T grad_unwrapped. Wrapped differences have slope 1 almost everywhere, so for
 * WRAPDIFF the Jacobian is the linear Poisson operator; SINCOS linearises at `wrapped`. */
ctdm_status ctdm_unwrap_dct_adjoint(ctdm_ctx *ctx, const float *wrapped, const float *grad_unwrapped,
                                    float *grad_wrapped, int batch, int M, unsigned flags);

/* Replaces cudaGoldsteinUnwrap2D (Goldstein.cu:5-50; residues :52-103, cuts :105-248, flood fill
 * :263-641) with a deterministic specification of the same steps, batched over `batch` H x W maps:
 *   residues  int8  +1/-1/0 (last row/column 0)
 *   cuts      uint8 1 on branch-cut pixels
 *   k2pi      int32 integer 2*pi count; INT32_MIN where the flood fill never reached the pixel
 *   unwrapped float wrapped + 2*pi*k2pi, 0 where unreached (as the reference leaves it)
 * Any output pointer except `unwrapped` may be NULL. Maps of at least 3 x 3 and below 2^30 pixels, at most
 * 65535 rows and 65535 maps per call; other sizes return CTDM_EINVAL. */
ctdm_status ctdm_goldstein(ctdm_ctx *ctx, const float *wrapped, int batch, int H, int W,
                           int8_t *residues, uint8_t *cuts, int32_t *k2pi, float *unwrapped);

/* ---------------------------------------------------------------- reconstruction (L4) */
/* Replaces CopyDataArray/cuCopyArray (Recon_kernel.cu:144-169) + cuFFT2Dshift + cuFFT2D(-1)
 * (HilbertTransform.cu:1909-1913): u = logamp + i*unwrapped, stored transposed, spatially
 * fftshifted, forward unnormalised 2-D DFT. A maps of M x M -> A spectra of M x M. */
ctdm_status ctdm_rytov_spectrum(ctdm_ctx *ctx, const float *unwrapped, const float *logamp, int A,
                                int M, ctdm_complex *spec);

/* Replaces EdwardSphere/fillEdwardSphere (Recon_kernel.cu:171-229) over all A angles:
 * F[v] += 4*pi*i*f_z*spec[a][p], hits[v] += 1 on the Ewald cap (ACCUMULATES: the caller zeroes
 * F and hits). angles: HOST array of A. M must equal V (as in the reference, ReconX = crop side). */
ctdm_status ctdm_ewald_scatter(ctdm_ctx *ctx, const ctdm_complex *spec, const ctdm_angle *angles,
                               int A, int M, int V, const ctdm_optics *optics, ctdm_complex *F,
                               float *hits);

/* Adjoint of ctdm_ewald_scatter (no reference counterpart): spec[a][p] = conj(4*pi*i*f_z) * F[v]
 * on the cap, 0 off the cap. Overwrites spec. */
ctdm_status ctdm_ewald_gather(ctdm_ctx *ctx, const ctdm_complex *F, const ctdm_angle *angles, int A,
                              int M, int V, const ctdm_optics *optics, ctdm_complex *spec);

/* Replaces initial_F_and_F2/Initial_F_F2 (Recon_kernel.cu:241-276): F /= hits where hits != 0,
 * F *= 1/dx, F_meas = F. F_meas may be NULL. */
ctdm_status ctdm_normalize_measured(ctdm_ctx *ctx, ctdm_complex *F, const float *hits,
                                    ctdm_complex *F_meas, int V, double dx);

/* Replaces cuFFT3D (FFT.cu:807-838) + scaleFFT3D (FFT.cu:900-916): in-place 3-D DFT of a V^3
 * volume; dir = -1 forward unnormalised, dir = +1 inverse scaled by 1/V^3 (no conjugation). */
ctdm_status ctdm_fft3d(ctdm_ctx *ctx, ctdm_complex *vol, int V, int dir);

/* Replaces the positivity loop of BatchRecon (HilbertTransform.cu:1940-1975): given the real-space
 * scattering potential f (the inverse 3-D DFT of the normalised measured spectrum; overwritten),
 * the measured spectrum F_meas and hit counts, runs `iters` iterations of
 *   f = modify(n); F = FFT3(f); F = F_meas where measured; f = iFFT3(F); n = est_n(f)
 * (calculate_n3D / Modify_F3D / check_F3D, Recon_kernel.cu:305-412) starting from n = est_n(f),
 * and writes n (complex RI; positivity Re n >= n_med, Im n >= 0) to `n`, octant-shifted unless
 * CTDM_POS_NO_SHIFT. iters = IterTime + 1 of the reference. hits may be NULL with
 * CTDM_DC_REFCOMPAT. CTDM_EXEC_PLAIN runs the unfused sequence. */
ctdm_status ctdm_positivity(ctdm_ctx *ctx, ctdm_complex *f, const ctdm_complex *F_meas,
                            const float *hits, ctdm_complex *n, int V, int iters,
                            const ctdm_optics *optics, unsigned flags);

/* Replaces MedianFilter_gpu (Recon_kernel.cu:65-119) as MedianFilter (:7-27) applies it: 3x3 median of
 * each H x W map of `in` into `out` (device, must not overlap); first and last rows and columns are
 * copied. The reference filters in place while neighbours are being overwritten (its result depends
 * on thread order); this filter reads only the unfiltered input. The nine values are gathered and
 * selected as the reference does (column-major 3x3 gather, fifth pick of its selection sort), so NaN
 * inputs resolve the same way. */
ctdm_status ctdm_median3x3(ctdm_ctx *ctx, const float *in, float *out, int batch, int H, int W);

/* Replaces checkMaps (Recon_kernel.cu:121-142), in place on n values (device; either may be NULL):
 * phase NaN/Inf -> 0; amplitude NaN/Inf -> 1, 0 -> 0.01, otherwise |A|. */
ctdm_status ctdm_check_maps(ctdm_ctx *ctx, float *phase, float *amp, size_t n);

/* Replaces BatchRecon (HilbertTransform.cu:1825-2047) from its own inputs: A unwrapped-phase maps and
 * A second maps (V x V each, device) and angles (HOST) -> Rytov spectra, Ewald gridding with hit
 * averaging, inverse 3-D DFT and `iters` positivity iterations; writes the octant-shifted complex
 * RI to `ri` (V^3, device). The second map is the log-amplitude, or with CTDM_MAPS_AMPLITUDE the
 * amplitude that the reference sanitises and logs itself (checkMaps, cuCopyArray,
 * Recon_kernel.cu:121-169). CTDM_MAPS_MEDIAN runs the reference's 3x3 median on both maps first
 * (BatchRecon, HilbertTransform.cu:1878). recon_flags: CTDM_DC_* | CTDM_POS_NO_SHIFT | CTDM_MAPS_* |
 * CTDM_EXEC_PLAIN (reference-structured kernel sequence). BatchRecon itself corresponds to
 * CTDM_MAPS_MEDIAN | CTDM_MAPS_AMPLITUDE. */
ctdm_status ctdm_recon_from_maps(ctdm_ctx *ctx, const float *unwrapped, const float *logamp,
                                 const ctdm_angle *angles, int A, int V, const ctdm_optics *optics,
                                 int iters, unsigned recon_flags, ctdm_complex *ri);

/* ---------------------------------------------------------------- pipeline */
/* One cell: A hologram pairs -> one V^3 RI volume (reference: per-frame loop
 * HilbertTransform.cu:785-1165 followed by BatchRecon, HilbertTransform.cu:1825-2047). */
typedef struct ctdm_cell {
    const void *holo;           /* A x H x W sample holograms (device): float32 pixels, or 8- / 16-bit
                                   unsigned pixels with CTDM_HOLO_U8 / CTDM_HOLO_U16 in recon_flags */
    const void *bg;             /* A x H x W backgrounds (device), same pixel format */
    const ctdm_angle *angles;   /* A angles (HOST) */
    int A, H, W;
    int V;                      /* volume side = crop side M */
    int iters;                  /* positivity iterations (reference IterTime + 1) */
    unsigned unwrap_flags;      /* CTDM_UNWRAP_* */
    unsigned recon_flags;       /* CTDM_DC_* | CTDM_POS_NO_SHIFT | CTDM_MAPS_MEDIAN | CTDM_RETRIEVE_ROWS |
                                   CTDM_HOLO_U8 or CTDM_HOLO_U16 */
    ctdm_optics optics;
    ctdm_crop crop;             /* crop.M == V; crop.radius < 0: locate from bg[0] (reference frame 1);
                                   ctdm_reconstruct writes back the crop it used */
    ctdm_complex *ri;           /* V^3 complex RI output (device) */
    float *wrapped_opt;         /* optional A x V x V outputs (device), may be NULL */
    float *unwrapped_opt;
    float *logamp_opt;
} ctdm_cell;

typedef struct ctdm_batch_opts {
    int mode;      /* 0 = plain (reference structure: per angle, unfused, host-synchronous);
                      1 = batched (all angles and cells per launch, fused passes) */
    int graph;     /* batched only: capture the launch sequence in a hipGraph and replay it */
    int autotune;  /* batched only: autotune kernel variants for the sizes present before capture */
    int lanes;     /* batched only: 1 = cells one after another on the context stream; 0 (default) or
                      2 = with two or more cells, cells alternate between the context stream and an
                      auxiliary stream forked from it and joined back before the call's work ends
                      (one cell's retrieval overlaps another's positivity loop; lane 2 has its own
                      scratch, so device memory for scratch doubles) */
    int reserved[4];
} ctdm_batch_opts;

/* Reconstructs every cell (no single reference counterpart: HilbertTransform()'s per-frame loop,
 * HilbertTransform.cu:951-1063, then BatchRecon). Batched mode locates each crop on the device, ahead
 * of the cell's work, and returns once those searches have finished (the cells' work is still running
 * on the context stream); cells[i].crop then holds the crop used. A located crop whose window leaves
 * the spectrum yields CTDM_EINVAL after the work was queued: the outputs are then undefined (the queued
 * kernels read the crop modulo the spectrum size and stay in bounds). Plain mode and row-wise
 * retrieval locate crops synchronously first. */
ctdm_status ctdm_reconstruct(ctdm_ctx *ctx, ctdm_cell *cells, int ncells,
                             const ctdm_batch_opts *opts);

/* Times the three fused volume passes (x with projection, y, z with data consistency) of each
 * listed volume side V with 8 and 16 pencils per workgroup on this device, inside the positivity
 * loop's own pass sequence, and caches the choice in the context: the faster one, except that y and
 * z take 8 pencils only when at least 3 % faster (their 8-pencil tiles write half lines). Then times the
 * loop iteration with the x or the y passes sweeping the z planes backwards (kept when at least 1 %
 * faster), the y pass with its alternative radix plan (kept when at least 1 % faster) and the gridding
 * pass with 8 and 16 pencils for 60 angles (16 when at least 3 % faster). The choices also serve contexts
 * created later in the process on the same GPU architecture. No reference counterpart. Every listed
 * side is re-timed, also one tuned before. Synchronises. `report` (HOST, may be NULL) receives a text
 * summary. */
ctdm_status ctdm_autotune(ctdm_ctx *ctx, const int *sizes, int nsizes, char *report,
                          size_t report_len);

/* The tuning table entry of volume side V (no reference counterpart): values[0..6] = pencils per workgroup
 * of the x, y, z+dc and gridding passes (8 or 16), the sweep direction of the x and y passes (0 forward,
 * 1 backwards over the z planes), and the radix plan of the y pass (0 the base factorisation, 1 the
 * alternative: 32 = 4x8, 64 = 4x4x4, 128 = 16x8, 192 = 16x12, 256 = 4x4x16). Pencil grouping and sweep
 * order compute bit-identical values; the two radix plans agree to FP32 rounding. ctdm_get_tuning reads the
 * entry (defaults, or choices tuned earlier in the process on the same architecture); ctdm_set_tuning
 * writes it and marks the side tuned. n must be 7. */
ctdm_status ctdm_get_tuning(ctdm_ctx *ctx, int V, int *values, int n);
ctdm_status ctdm_set_tuning(ctdm_ctx *ctx, int V, const int *values, int n);

/* Average duration [ms] of one launch of a fused volume pass (pass 0: x with projection, 1: y,
 * 2: z with data consistency; 4: a plain streaming copy of the V^3 complex64 volume, the bandwidth
 * yardstick) at volume side V with `pencils` pencils per workgroup (8 or 16;
 * <= 0 = the tuned choice), timed with hipEvents on the context stream over `reps` launches on the
 * context's volume buffer (no reference counterpart: the measuring primitive of ctdm_autotune).
 * *bytes receives the algorithmic bytes of one launch. Synchronises. */
ctdm_status ctdm_time_pass(ctdm_ctx *ctx, int pass, int V, int pencils, int reps, double *avg_ms,
                           double *bytes);

/* Runs `iters` iterations of the fused positivity loop's pass sequence (x with projection, forward y,
 * z with data consistency, inverse y) on the context's volume of side V, each launch carrying its own
 * start and end events (hipExtLaunchKernel), and returns the average duration [ms] of each pass in that
 * sequence
 * (avg_ms[0] x, [1] y, [2] z + data consistency; host arrays of 3) and its algorithmic bytes per
 * launch (bytes[3]). Uses the data-consistency list of the last reconstruction at this V, when there is
 * one. No reference counterpart: the per-kernel measurement of the loop as it runs. Synchronises. */
ctdm_status ctdm_time_loop(ctdm_ctx *ctx, int V, int iters, double *avg_ms, double *bytes);

/* ---------------------------------------------------------------- on-disk formats (host only) */
/* The reference's per-frame .phimap file (outputImg / openImage, HilbertTransform.cu:2437-2477), written
 * as the reference writes it: uint8 status, float64 angle x, float64 angle y, int32 sizeX, int32 sizeY,
 * then sizeX*sizeY (phase, amplitude) float32 pairs (microImg), little endian, no padding. All pointers
 * HOST. status 0 marks a frame the reference's quality check rejected (RefreshStack drops it). */
ctdm_status ctdm_phimap_write(const char *path, int status, double ang_x, double ang_y, int sizeX, int sizeY,
                              const float *phase, const float *amp);
/* Reads a .phimap. Header outputs may be NULL; with phase = amp = NULL only the header is read,
 * otherwise `capacity` (elements of each array) must hold sizeX*sizeY. */
ctdm_status ctdm_phimap_read(const char *path, int *status, double *ang_x, double *ang_y, int *sizeX,
                             int *sizeY, float *phase, float *amp, size_t capacity);
/* The reconstructed volume as BatchRecon exports it (HilbertTransform.cu:2013-2041, exportComplex /
 * exportRAW_F :3457-3487): part 0 = raw complex64, 1 = real floats, 2 = imaginary floats. HOST arrays
 * of n elements; reading part 1 or 2 fills only that component. */
ctdm_status ctdm_rimap_write(const char *path, const ctdm_complex *ri, size_t n, int part);
ctdm_status ctdm_rimap_read(const char *path, ctdm_complex *ri, size_t n, int part);

/* ---------------------------------------------------------------- differentiable fit */
/* Projected-gradient fit of measured Rytov spectra (no reference counterpart; an alternative to the
 * data-consistency loop of ctdm_positivity). Minimises
 *     J(f) = 1/2 sum_a sum_p |FFT3(f)[v_a(p)] - m_a(p)|^2,   m_a(p) = 4*pi*i*f_z * spec_a(p) / dx,
 * over the scattering potential f (v_a(p) and f_z as in ctdm_ewald_scatter). Each of `iters` steps
 * recomputes F = FFT3(f), takes the hit-preconditioned gradient step F -= step * R / hits with
 * R = G^T (G F - m) (G the gather, G^T the scatter), returns to real space and projects onto the
 * positivity set (Re n >= n_med, Im n >= 0). G picks one voxel per (angle, pencil), so G^T G is the
 * diagonal of hit counts and the step equals F += step * (m_bar - F) on measured voxels, m_bar the
 * hit-averaged scatter of m; m_bar is formed once per call. step = 1 reproduces ctdm_positivity fed
 * with the averaged measurement. spec: A x V x V Rytov spectra (ctdm_rytov_spectrum). f: start potential,
 * used as workspace (overwritten). n: est_n of the result, octant-shifted unless CTDM_POS_NO_SHIFT. */
ctdm_status ctdm_fit_pgd(ctdm_ctx *ctx, ctdm_complex *f, const ctdm_complex *spec,
                         const ctdm_angle *angles, int A, int V, const ctdm_optics *optics,
                         int iters, double step, ctdm_complex *n, unsigned flags);

/* ctdm_fit_pgd that also reports its objective (no reference counterpart): objective (HOST, iters values)
 * receives J before each step, J_k = 1/2 sum_a sum_p |FFT3(P f_k)[v_a(p)] - m_a(p)|^2 with P f_k the
 * positivity-projected iterate the step starts from (f_0 the given start). The z pass of each step sums the
 * measured voxels' h |F - m-bar|^2 / 2 and the gridding kernel the constant sum of (sum |m|^2 - h |m-bar|^2) / 2
 * (h: hits of the voxel), which together equal J exactly. Same result n as ctdm_fit_pgd. Synchronises. */
ctdm_status ctdm_fit_pgd_trace(ctdm_ctx *ctx, ctdm_complex *f, const ctdm_complex *spec,
                               const ctdm_angle *angles, int A, int V, const ctdm_optics *optics,
                               int iters, double step, ctdm_complex *n, unsigned flags, double *objective);

/* ---------------------------------------------------------------- synthetic inputs (host) */
/* Seeded synthetic phantoms and holograms (host memory; no reference counterpart, the
 * reference reads BMP files). See the generator source for the models. */
typedef struct ctdm_synth_params {
    int H;              /* hologram side (1024) */
    double pixel;       /* sample-plane hologram pixel [m] (0.1e-6) */
    double wavelength;  /* [m] */
    double n_med;
    double na_ill;      /* illumination NA of the angle cone */
    double n_cell, n_nucleoli;
    double kappa_cell;  /* imaginary RI of the cell (weak absorption) */
    double semi_axes[3];/* ellipsoid semi-axes [m] */
    int n_nuc;          /* number of nucleoli */
    double r_nuc;       /* nucleolus radius [m] */
    double noise;       /* Gaussian intensity noise sigma */
    double carrier[2];  /* carrier frequency [cycles / pixel] along x and y */
    double field_noise; /* sigma of complex Gaussian noise added to the sample field A exp(i phi) */
    int model;          /* 1 = first-order Rytov fields of the phantom (default), 0 = straight-ray projections */
} ctdm_synth_params;

void ctdm_synth_default_params(ctdm_synth_params *p);
/* A angles on a cone (n_med sin theta = na_ill, azimuth 2*pi*a/A; ang_y is the tilt along image y, as the
 * reference's transposed gridding reads it) -> angles_host[A]; holo_host/bg_host: A x H x H float32 each
 * (host), H a power of two. */
ctdm_status ctdm_synth_cell(uint64_t seed, int A, const ctdm_synth_params *p, float *holo_host,
                            float *bg_host, ctdm_angle *angles_host);
/* The refractive index (n + i kappa) of the phantom ctdm_synth_cell images with the same seed and
 * parameters, sampled at the voxel centres of a V^3 reconstruction of pitch dx [m] in the pipeline's
 * output orientation (volume x = image y, volume y = image x, voxel (V/2, V/2, V/2) at the origin);
 * n_host: V^3 values, idx = x + V*y + V*V*z (host). */
ctdm_status ctdm_synth_phantom(uint64_t seed, const ctdm_synth_params *p, int V, double dx,
                               ctdm_complex *n_host);
/* Single off-axis hologram (sphere phase of radius 80 px, peak 12 rad, plus a 2nd-order
 * background) and a background, H x H each (host). */
ctdm_status ctdm_synth_hologram(uint64_t seed, int H, float *holo_host, float *bg_host);
/* `batch` wrapped phase maps of H x W: sum of 8 Gaussians (|peak| <= 40 rad) + U(-0.8pi, 0.8pi)
 * noise, wrapped to (-pi, pi] (host). */
ctdm_status ctdm_synth_wrapped(uint64_t seed, int batch, int H, int W, float *out_host);

#ifdef __cplusplus
}
#endif
#endif /* CTDM_H */
