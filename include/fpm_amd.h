/*
 * fpm_amd.h -- C-ABI of the MI355X (gfx950) fused Fourier ptychographic
 * microscopy (FPM) reconstruction library, libfpm_amd.so.
 *
 * Every array is column-major exactly as MATLAB lays it out: element
 * (row i, column j) of an R x C array sits at j*R + i.  Complex values are
 * interleaved {re, im} float pairs.  All entry points return 0 on success or
 * a negative FPM_E* code; fpm_last_error() then returns a thread-local text.
 * No entry point mutates its inputs.  Outputs are caller-allocated.
 *
 * Reference interfaces replaced (paths relative to the reference repository
 * THUHoloLab/Fused-Fourier-Ptycho):
 *   fpm_pure_v3          <- mexFunction of cuFPM_v3/cuFPM_pure_v3.cu:44-228
 *                           (argument list prhs[0..8], outputs plhs[0..1])
 *   fpm_fused_gradient   <- mexFunction of benchmark/cuFPM/fullyfusedFPM.cu:21-162
 *                           (w1, w2, Y, ledIdx, pratio) -> (dldw1, dldw2), with
 *                           the v3 loss of cuFPM_v3/cuda/kernels.cuh:197-275
 *   fpm_led_offsets      <- codes_v2/func_ware/init_environment_rgb.m:33-134 and
 *                           codes_v2/main_run_FPM.m:23,27-31,53-55
 *   fpm_led_offsets_at   <- the same with the field-area offsets objdx, objdy
 *                           of init_environment_rgb.m:26-29,83-84
 *   fpm_prepare_amplitude<- codes_v2/main_run_FPM.m:40-45
 *   fpm_forward_intensity<- the forward half of benchmark/func_ware/fpm_forward.m:25-35
 *                           (|IFFT(crop(FFT O) . P)|, the image model)
 *   fpm_fft2             <- the cufftExecC2C call sites of cuFPM_v3/cuFPM_pure_v3.cu:118-187
 *                           (unnormalised 2D C2C transforms)
 *   fpm_optimizer_step   <- optimizer_*.step of codes/func_ware/optimizers/ (adam, RMSprop,
 *                           sgd) and codes_v2/func_ware/optimizers/optimizer_yogi.m, the
 *                           optimizers that drive fullyfusedFPM in codes/main_run_FPM.m:136-151
 *   fpm_pupil_constraint <- the pupil amplitude constraint of codes/main_run_FPM.m:153-156
 */
#ifndef FPM_AMD_H
#define FPM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FPM_AMD_ABI_VERSION 3

/* status codes */
#define FPM_OK              0
#define FPM_E_INVALID      -1  /* bad argument (null pointer, bad shape, batch <= 0 ...) */
#define FPM_E_UNSUPPORTED  -2  /* shape outside what the kernels handle (see below) */
#define FPM_E_HIP          -3  /* a HIP runtime call failed (text says which) */
#define FPM_E_NOMEM        -4  /* device allocation failed */
#define FPM_E_NODEVICE     -5  /* no usable gfx950 device */

/* shape rules: every image side (rows and columns of the object, pupil and
 * images) is even and in [16, 16384], and the object is at least as large as
 * the pupil in each dimension (the reference takes any even side: cuFFT plans
 * of any length, README.md:9, cuFPM_pure_v3.cu:118-126).  When all four sides
 * are powers of two the solve runs the fused kernels; otherwise every call
 * runs the generic path: FP32 mixed-radix transforms (radices 2, 3, 4, 5, 7
 * as butterflies, any other prime factor as a direct sum) and unfused
 * kernels, with the owner-computes stitch.  On the generic path
 * fft_engine = FPM_FFT_SPLIT_F16 and stitch = FPM_STITCH_ATOMIC return
 * FPM_E_UNSUPPORTED (fpm_tiles_solve runs such tiles with FP32 transforms and
 * applies the split engine to its power-of-two tiles). */

typedef struct { float re, im; } fpm_c32;

/* type of the image stack Y */
#define FPM_Y_F32_AMPLITUDE 0  /* n x n x L single amplitudes: exactly what v3 takes */
#define FPM_Y_U16_RAW       1  /* n x n x L uint16 raw counts; the library applies
                                  Y = sqrt((I - min)/(max - min)) in float, as
                                  codes_v2/main_run_FPM.m:43-45 does in single */

/* how the object-spectrum gradient is stitched from the per-LED back-projections */
#define FPM_STITCH_ATOMIC 0  /* FP32 atomics into the spectrum (v3 kernels.cuh:115-116) */
#define FPM_STITCH_OWNER  1  /* owner-computes gather in fixed LED order: deterministic;
                                the default (faster on MI355X: it is formed inside the
                                first column pass of the gradient's inverse transform) */

/* launches of fpm_tiles_solve (fpm_opts.tile_launch) */
#define FPM_TILES_GROUPED  0  /* tiles of one size and image type share launches */
#define FPM_TILES_PER_TILE 1  /* one launch sequence per tile */

/* arithmetic of the radix-16 DFT stages of every transform of a call */
#define FPM_FFT_F32       0  /* FP32 butterflies on the vector ALUs (the default) */
#define FPM_FFT_SPLIT_F16 1  /* radix-16 DFTs as FP16 matrix-core products (MFMA) with
                                split hi + lo operands, per-butterfly power-of-two
                                scaling and FP32 accumulation; twiddles, small-radix
                                passes and all non-FFT arithmetic stay FP32 */

/* Per-stage device time of one call, measured with HIP events recorded on the
 * call's stream around each stage (opts.stage_times; costs a few us per step). */
enum {
    FPM_STAGE_SPECTRUM_INIT = 0, /* DFT of the initial object (contiguous + column passes) */
    FPM_STAGE_CROP,              /* crop x pupil (kernels.cuh:19-65) */
    FPM_STAGE_IFFT_N_COLS_LO,    /* batched inverse DFT, column pass(es) */
    FPM_STAGE_IFFT_N_CONTIG_LO,  /* batched inverse DFT, contiguous pass */
    FPM_STAGE_STENCIL,           /* loss stencil (kernels.cuh:197-275) */
    FPM_STAGE_FFT_N_CONTIG_LO,   /* batched forward DFT, contiguous pass */
    FPM_STAGE_FFT_N_COLS_LO,     /* batched forward DFT, column pass(es) */
    FPM_STAGE_ZERO,              /* clearing the spectrum gradient (atomic stitch) */
    FPM_STAGE_STITCH_OWNER,      /* owner-computes stitch */
    FPM_STAGE_BACKPROJ,          /* back-projection (+ atomics) + pupil update */
    FPM_STAGE_IFFT_N_COLS_HI,    /* inverse DFT of the spectrum gradient, column passes */
    FPM_STAGE_FUSED_UPDATE,      /* contiguous inverse pass + RMSprop of O + contiguous forward pass */
    FPM_STAGE_FFT_N_COLS_HI,     /* forward DFT of the new object, column passes */
    FPM_STAGE_COUNT
};

typedef struct fpm_stage_times {
    double  ms[FPM_STAGE_COUNT];      /* summed device time per stage */
    int64_t count[FPM_STAGE_COUNT];   /* number of stage executions */
    int32_t launches[FPM_STAGE_COUNT];/* kernel launches per execution */
} fpm_stage_times;

typedef struct {
    int32_t struct_size;          /* sizeof(fpm_opts); lets the ABI grow */
    int32_t stitch;               /* FPM_STITCH_*; default OWNER */
    int32_t pointers_are_device;  /* 1: every array argument is a device pointer */
    int32_t verbose;              /* 1: print the two parameter lines v3 prints */
    void*   stream;               /* hipStream_t to run on; NULL = a library stream
                                     that starts after the work queued on the
                                     null stream when the call is made */
    double* loss_per_epoch;       /* optional host array [epochs]: FP64 sum over the
                                     epoch's minibatches of sum|Dx e| + |Dy e| */
    float   pupil_clamp;          /* pupil amplitude band (v3 hard-codes 0.2);
                                     <= 0 selects 0.2 */
    fpm_stage_times* stage_times; /* optional: filled with per-stage device time */
    int32_t graph;                /* 1: each epoch's launches are captured once as a
                                     hipGraph and replayed (at most four distinct
                                     epoch graphs per call); ignored when stage_times
                                     is set (per-stage events need eager launches) */
    int32_t fft_engine;           /* FPM_FFT_*; default FPM_FFT_F32 */
    int32_t lr_halve_from;        /* lr halves after every epoch with index >= this;
                                     0 selects 13, the reference's schedule
                                     (cuFPM_pure_v3.cu:204-206: index > 12) */
    int32_t tile_launch;          /* fpm_tiles_solve: FPM_TILES_GROUPED (default) or
                                     FPM_TILES_PER_TILE */
    int32_t reserved;
    double  stop_loss;            /* > 0: fpm_pure_v3 stops after the first epoch whose
                                     loss (as loss_per_epoch defines it) is <= stop_loss;
                                     the host reads each epoch's loss (one stream
                                     synchronisation per epoch); loss_per_epoch entries
                                     past the last epoch run are set to NaN */
    int32_t* epochs_run;          /* optional host int: epochs fpm_pure_v3 ran */
    uint8_t* stencil_signs;       /* optional, fpm_fused_gradient only (host or device
                                     like the other arrays): n_rows x n_cols x L bytes
                                     receiving the loss stencil's sign decisions, pixel
                                     (i, j) of image l at l n_r n_c + j n_r + i:
                                     bit 0 = sign bit of e(i, j+1) - e(i, j), bit 1 =
                                     sign bit of e(i+1, j) - e(i, j), cyclic, with
                                     e = |z|/(n_r n_c r^2) - Y.  These are the two
                                     copysign decisions per pixel that
                                     kernels.cuh:234-253 combines; given to an FP64
                                     computation they separate its distance from the
                                     GPU's into arithmetic and near-tie flips.  No
                                     counterpart in the reference. (ABI 3) */
} fpm_opts;

/* Fills *o with the defaults (owner-computes stitch, host pointers, silent,
 * library stream, no loss, clamp 0.2, eager launches, FP32 transforms). */
void fpm_opts_init(fpm_opts* o);

/* The drop-in solve: [O, P] = cuFPM_pure_v3(O0, P0, Y, ledIdx, A, pratio,
 * numEpochs, batch, [lr beta eps]).
 *   O0      N_rows x N_cols complex single, SPATIAL-domain object
 *   P0      n_rows x n_cols complex single pupil
 *   Y       n_rows x n_cols x L images (y_dtype says float amplitude or uint16 raw)
 *   led_idx int32 2 x L, column l = {kl (1-based column start), kt (1-based row
 *           start)} of LED l's window in the centred object spectrum
 *   A       n_rows x n_cols single pupil mask
 *   pratio, epochs, batch: doubles truncated to int as the mex does
 *   opt     {learning rate, beta, eps}, cast to float
 *   O_out, P_out: results (may alias O0 / P0)
 * opts may be NULL (v3 semantics, host pointers). */
int fpm_pure_v3(const fpm_c32* O0, int64_t N_rows, int64_t N_cols,
                const fpm_c32* P0, int64_t n_rows, int64_t n_cols,
                const void* Y, int32_t y_dtype, int64_t L,
                const int32_t* led_idx, const float* A,
                double pratio, double epochs, double batch, const double opt[3],
                fpm_c32* O_out, fpm_c32* P_out, const fpm_opts* opts);

/* One fused gradient over all L images (one minibatch):
 *   G_O = IDFT_N(scatter(R . conj(P) / (N_rows N_cols))),  G_P = sum_l R_l . conj(S_l)
 * where R_l = fftshift(DFT_n(u_l)) and u_l is the feature-domain loss
 * gradient of v3 (kernels.cuh:197-275).  Equals r^2 grad_O Loss and
 * n^2 r^4 grad_P Loss for N = n r.  loss_out (optional, host) receives the
 * FP64 loss sum|Dx e| + |Dy e| at (O, P). */
int fpm_fused_gradient(const fpm_c32* O, int64_t N_rows, int64_t N_cols,
                       const fpm_c32* P, int64_t n_rows, int64_t n_cols,
                       const void* Y, int32_t y_dtype, int64_t L,
                       const int32_t* led_idx, double pratio,
                       fpm_c32* G_O, fpm_c32* G_P, double* loss_out,
                       const fpm_opts* opts);

/* fpm_fused_gradient for callers that cannot allocate device memory.  No
 * counterpart in the reference: it serves code that must not call hipMalloc
 * (work under HIP graph capture, serving loops, optimizer loops that forbid
 * allocation, e.g. a gradient-driven solve driving this entry every step).
 * Every array is a DEVICE pointer: O, P, Y (single amplitudes, as v3 takes),
 * led_idx, G_O, G_P and the optional loss (one device double, cleared by a
 * fill kernel, then the FP64 loss sum|Dx e| + |Dy e|).  Power-of-two sides,
 * FP32 transforms, owner-computes stitch: the same kernels and bits as
 * fpm_fused_gradient with those options.  Queued on `stream` (NULL: the null
 * stream) and not waited for.  The working buffers (two N_rows x N_cols
 * spectra, two sets of L per-LED pages, a pupil copy: 16 (N_rows N_cols +
 * L n_rows n_cols) + 8 n_rows n_cols bytes, each rounded up to 256) live in a
 * static device array of FPM_NOALLOC_SCRATCH_BYTES (64 MiB), one per GPU,
 * held until the process ends: e.g. a 1024^2 object with 26 images of 256^2
 * fits; larger calls return FPM_E_INVALID.  Calls that overlap on one GPU
 * (several host threads or streams) race on that array. */
#define FPM_NOALLOC_SCRATCH_BYTES (64u << 20)
int fpm_fused_gradient_noalloc(const fpm_c32* O, int64_t N_rows, int64_t N_cols,
                               const fpm_c32* P, int64_t n_rows, int64_t n_cols,
                               const float* Y, int64_t L, const int32_t* led_idx, double pratio,
                               fpm_c32* G_O, fpm_c32* G_P, double* loss, void* stream);

/* Optimizers of the gradient-driven solve (fpm_fused_gradient in a loop, as
 * codes/main_run_FPM.m:106-160 drives fullyfusedFPM): one step on DEVICE
 * arrays of `count` complex elements, in place on param, with the
 * reference's MATLAB arithmetic in single precision:
 *   FPM_OPT_ADAM     m1 = b1 m1 + (1-b1) g;  m2 = b2 m2 + (1-b2)|g|^2;
 *                    param -= lr (bias m1) / (sqrt(m2) + 1e-5),
 *                    bias = sqrt(1 - b2^iter) / (1 - b1^iter)   (optimizer_adam.m:25-38)
 *   FPM_OPT_YOGI     m1 as Adam; m2 -= (1-b2) sign(m2 - |g|^2) |g|^2;
 *                    param -= lr m1 / (sqrt(m2) + 1e-5)          (optimizer_yogi.m:25-36)
 *   FPM_OPT_RMSPROP  s = b1 + (1-b1)|g|^2 (no running average, as written);
 *                    param -= lr g / (sqrt(s) + 1e-5)            (optimizer_RMSprop.m:28-46,
 *                    centered = false, mum = 0 as every caller constructs it)
 *   FPM_OPT_SGD      param -= lr g  (optimizer_sgd.m:16-21 subtracts an
 *                    undefined `update`; the plain gradient step is meant)
 * mom1 (complex) and mom2 (float) hold the moments (zero before the first
 * step; unused by RMSprop and SGD, may be NULL there); iter counts from 1
 * (Adam's bias correction).  Scalars are rounded to single as MATLAB does
 * when a double meets a single array ((1 - b) is rounded once). */
#define FPM_OPT_SGD     0
#define FPM_OPT_ADAM    1
#define FPM_OPT_YOGI    2
#define FPM_OPT_RMSPROP 3
int fpm_optimizer_step(int32_t kind, fpm_c32* param, const fpm_c32* grad, fpm_c32* mom1,
                       float* mom2, int64_t count, double lr, double beta1, double beta2,
                       int64_t iter, void* stream);

/* The gradient-driven solve's pupil constraint (codes/main_run_FPM.m:153-156),
 * in place on DEVICE arrays:
 *   P = max(min(|P|, A + ac), A - ac) . sign(P) . A,   sign(P) = P / |P| (0 at 0).
 * (fpm_pure_v3 applies its own clamp inside the pupil update.) */
int fpm_pupil_constraint(fpm_c32* P, const float* A, int64_t count,
                         double amplitude_constraint, void* stream);

/* Optics set-up: LED illumination wave-vectors -> integer window offsets,
 * computed in FP64 (init_environment_rgb.m:33-134, main_run_FPM.m:23,53-55).
 *   wavelength_um, na, pixel_um (camera pixel / magnification, in um),
 *   led_pitch_um, led_height_um, led_grid (odd, e.g. 19), n (image side),
 *   pratio (upsampling) ; spiral = 1 orders LEDs in the centre-out spiral of
 *   init_environment_rgb.m:110-134, 0 keeps row-major grid order.
 * Outputs: led_idx (int32 2 x L, {kl, kt}), grid_index (int32 L, 1-based grid
 * number k(i,j) = j + (i-1)*led_grid of each entry, may be NULL),
 * na_illum (double L, may be NULL).  L = led_grid^2.
 * Returns FPM_E_INVALID if an argmin is not unique (a tie). */
int fpm_led_offsets(double wavelength_um, double na, double pixel_um,
                    double led_pitch_um, double led_height_um,
                    int32_t led_grid, int64_t n, int32_t pratio, int32_t spiral,
                    int32_t* led_idx, int32_t* grid_index, double* na_illum);

/* fpm_led_offsets for the field area at (objdx_um, objdy_um) of the sample
 * plane: the LED lateral offsets (-v D, -h D) become (-v D + objdx, -h D + objdy)
 * as in codes_v2/func_ware/init_environment_rgb.m:26-29,83-84 (the reference
 * fixes objdx = objdy = 0 there, i.e. fpm_led_offsets).  Gives each
 * field-of-view tile of fpm_tiles_solve its own illumination angles. */
int fpm_led_offsets_at(double objdx_um, double objdy_um, double wavelength_um, double na,
                       double pixel_um, double led_pitch_um, double led_height_um,
                       int32_t led_grid, int64_t n, int32_t pratio, int32_t spiral,
                       int32_t* led_idx, int32_t* grid_index, double* na_illum);

/* uint16 raw stack -> float amplitudes sqrt((I - min)/(max - min)) with min and
 * max taken over the whole stack (main_run_FPM.m:43-45).  count = n*n*L.
 * Pointers are host or device per pointers_are_device; min/max outputs may be NULL. */
int fpm_prepare_amplitude(const uint16_t* raw, int64_t count, float* amp_out,
                          int32_t pointers_are_device, void* stream,
                          uint16_t* min_out, uint16_t* max_out);

/* Image model of the solver: amp_l = |IDFT_n(ifftshift(crop_l(DFT_N(O)) . P))|
 * / (n_rows n_cols r^2) with r = (int)pratio -- the amplitude v3 compares with
 * Y (kernels.cuh:223,233) -- for l in [0, L), written as n_rows x n_cols x L
 * float.  With squared = 1 the intensities amp^2 are written instead.  The
 * crop follows the same led_idx convention as fpm_pure_v3. */
int fpm_forward_intensity(const fpm_c32* O, int64_t N_rows, int64_t N_cols,
                          const fpm_c32* P, int64_t n_rows, int64_t n_cols,
                          int64_t L, const int32_t* led_idx, double pratio,
                          int32_t squared, float* out,
                          int32_t pointers_are_device, void* stream);

/* Unnormalised 2D C2C transform of `batch` column-major rows x cols pages,
 * out-of-place or in place (out == in), natural order in and out.
 * direction -1 = forward (exp(-2 pi i ...)), +1 = inverse (no 1/N scaling). */
int fpm_fft2(const fpm_c32* in, fpm_c32* out, int64_t rows, int64_t cols,
             int64_t batch, int32_t direction, int32_t pointers_are_device,
             void* stream);

/* fpm_fft2 with the arithmetic of the radix-16 stages chosen by engine
 * (FPM_FFT_F32 or FPM_FFT_SPLIT_F16).  No counterpart in the reference: it
 * exposes the split-precision transform for accuracy checks against FP64. */
int fpm_fft2_engine(const fpm_c32* in, fpm_c32* out, int64_t rows, int64_t cols,
                    int64_t batch, int32_t direction, int32_t engine,
                    int32_t pointers_are_device, void* stream);

/* ------------------------------------------------------------------------
 * Multi-tile solve.  No counterpart in the reference (one field of view per
 * mex call): many field-of-view tiles, each with its own object, pupil,
 * images and LED windows (its own illumination angles), reconstructed
 * jointly with the update of fpm_pure_v3.  Every tile has the same number of
 * images L and uses the same pratio, epochs, batch and RMSprop constants;
 * tile sides are powers of two (the transforms' rule).  All tile arrays are
 * DEVICE pointers (opts->pointers_are_device is ignored).
 * With opts->tile_launch = FPM_TILES_GROUPED (the default) the tiles of one
 * size and image type run through shared launches: every stage of a
 * minibatch step takes all of them at once (the library holds their working
 * arrays back to back and reads each tile's images, LED windows and mask
 * through a device table), so an epoch costs the launches of one solve per
 * tile size instead of one per tile.  Tiles that cannot be grouped (atomic
 * stitch, batch > 64, objects narrower than a fused-update workgroup) and
 * every tile under FPM_TILES_PER_TILE run their own launch sequences.
 * Each epoch of all launch sequences is captured once as one hipGraph whose
 * sequences are spread over `streams` concurrent streams, largest first
 * (longest-processing-time assignment); streams = 0 autotunes: the first
 * epoch runs the sequences one after another and times each, and these
 * measured times order and assign them from then on; the next epochs run
 * with 1, 2, 4 and 8 streams in turn (at most one per sequence), each timed,
 * and the fastest count runs the rest.  Grouping and the stream
 * count change only the launches, never the arithmetic: every tile's result
 * is the same bits as its own fpm_pure_v3 call.
 * ------------------------------------------------------------------------ */
typedef struct {
    const fpm_c32* O0;       /* N x N spatial object */
    fpm_c32* O_out;          /* N x N result (may alias O0) */
    int64_t N;
    const fpm_c32* P0;       /* n x n pupil */
    fpm_c32* P_out;          /* n x n result (may alias P0) */
    int64_t n;
    const void* Y;           /* n x n x L images */
    int32_t y_dtype;         /* FPM_Y_* */
    const int32_t* led_idx;  /* int32 2 x L {kl, kt} of this tile */
    const float* A;          /* n x n pupil mask */
    double* loss_per_epoch;  /* optional HOST array [epochs] */
} fpm_tile;

typedef struct {
    int32_t streams;         /* stream count of the epochs after tuning */
    int32_t tuned;           /* candidates timed (0: streams was given) */
    int32_t cand[4];         /* candidate stream counts */
    double  cand_ms[4];      /* measured epoch time of each candidate */
    int32_t sequences;       /* launch sequences per epoch (groups + single tiles) */
    int32_t grouped_tiles;   /* tiles that ran in shared launches */
} fpm_tiles_report;

int fpm_tiles_solve(const fpm_tile* tiles, int32_t ntiles, int64_t L, double pratio,
                    double epochs, double batch, const double opt[3], int32_t streams,
                    const fpm_opts* opts, fpm_tiles_report* report);

/* ------------------------------------------------------------------------
 * Learned warm start.  No counterpart in the reference (its solve starts from
 * the upsampled centre image, codes_v2/main_run_FPM.m:48): a U-Net of 3x3
 * convolutions maps low-resolution features of the image stack to the
 * initial amplitude and phase of the object; the plain solve then starts from
 * it and finishes the reconstruction.  Training and inference run on the
 * library's own BF16 matrix-core kernels with FP32 accumulation.
 * Images here are planes in row-major order (element (y, x) at y*W + x);
 * feats, targets and out are DEVICE pointers whose producers have finished;
 * every call synchronises the work it queued before returning.
 * ------------------------------------------------------------------------ */
typedef struct fpm_cnn fpm_cnn;
typedef struct {
    int32_t levels;        /* U-Net levels (resolutions n, n/2, ...), 1..6 */
    int32_t base_channels; /* channels of level 0; level l has base * 2^l (<= 256) */
    int32_t in_channels;   /* feature planes per sample */
    int32_t n;             /* feature side in pixels */
    int32_t up;            /* output side = up * n; 2 up^2 a multiple of 32 */
    int32_t max_batch;     /* largest batch of one call (sizes the workspaces; <= 64 with a pupil head) */
    int32_t pupil_terms;   /* K coefficients of the pupil phase predicted from the mean of the bottom
                              level's features (64 hidden, ReLU); 0: no pupil head; <= 16 */
    float pupil_weight;    /* weight of the coefficients' mean squared error in the training loss */
} fpm_cnn_desc;

/* Network with He-normal weights from `seed`, zero biases, Adam state zero. */
int fpm_cnn_create(const fpm_cnn_desc* desc, uint64_t seed, fpm_cnn** out);
void fpm_cnn_destroy(fpm_cnn* net);
/* Number of FP32 parameters: per conv (encoder levels, decoder levels, head)
 * W (Cout x 9 x Cin, Cin of the first conv rounded up to 32) then b (Cout);
 * then, with a pupil head, W1 (64 x Cb), b1 (64), W2 (K x 64), b2 (K), Cb the
 * bottom level's channel count. */
int64_t fpm_cnn_num_params(const fpm_cnn* net);
int fpm_cnn_get_params(fpm_cnn* net, float* host_out);
/* Sets the parameters (host array) and clears the optimiser state. */
int fpm_cnn_set_params(fpm_cnn* net, const float* host_in);
/* feats: batch x in_channels x n x n; out: batch x 2 x (up n) x (up n)
 * (plane 0 amplitude, plane 1 phase); pupil_out: batch x pupil_terms
 * coefficients (NULL: not written; must be NULL without a pupil head) */
int fpm_cnn_forward(fpm_cnn* net, const float* feats, int64_t batch, float* out, float* pupil_out);
/* Mean squared error against targets (batch x 2 x (up n)^2), plus
 * pupil_weight x the mean squared error of the coefficients against
 * pupil_targets (batch x pupil_terms; required with a pupil head, NULL
 * without), and its gradient with respect to every parameter (host_grad may
 * be NULL). */
int fpm_cnn_gradient(fpm_cnn* net, const float* feats, const float* targets, const float* pupil_targets,
                     int64_t batch, double* loss_out, float* host_grad);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss. */
int fpm_cnn_train_step(fpm_cnn* net, const float* feats, const float* targets, const float* pupil_targets,
                       int64_t batch, float lr, double* loss_out);

/* Pupil estimate of the warm start (no counterpart in the reference): under
 * tilted illumination k_i the pupil phase's gradient at k_i shifts the
 * brightfield image, so the brightfield images' shifts against the centre
 * image measure the pupil's aberrations (their Zernike coefficients follow by
 * a linear map fitted on training data, fpm_amd.warmstart.ParallaxPupil).
 * For sample b and j < count: shifts[(b count + j) 2 + {0, 1}] = (dy, dx),
 * the global shift with plane which[j](y, x) ~ plane ref(y - dy, x - dx), by
 * `iters` Lucas-Kanade iterations on the mean/std-normalised planes
 * (bilinear warp with clamped coordinates, interior pixels, FP64 sums).
 * planes: DEVICE, batch x planes_per_sample row-major rows x cols FP32 planes
 * (3 <= rows, 3 <= cols <= 2730);
 * which: HOST, count (<= 64) plane indices; shifts: DEVICE; stream: a
 * hipStream_t or NULL; synchronises it before returning. */
int fpm_parallax_shifts(const float* planes, int64_t batch, int64_t planes_per_sample, int64_t rows, int64_t cols,
                        int32_t ref, const int32_t* which, int32_t count, int32_t iters, float* shifts,
                        void* stream);

/* Thread-local text of the last error ("" if none). */
const char* fpm_last_error(void);

/* ABI version (FPM_AMD_ABI_VERSION). */
int32_t fpm_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FPM_AMD_H */
