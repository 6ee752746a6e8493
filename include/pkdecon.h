/*
 * pkdecon.h -- C-ABI of the MI355X (gfx950) deconvolution engine.
 *
 * Drop-in boundary for PetaKit5D's RL / OMW deconvolution hot path.  Each
 * entry point names the reference interface it replaces (paths relative to
 * the PetaKit5D tree).  A MATLAB MEX gateway, a ctypes loader or a C/C++
 * program binds these symbols; no C++ or torch types cross the boundary.
 *
 * Conventions
 *  - Volumes are MATLAB column-major: dims[0] (n1) is the contiguous axis.
 *  - 64-bit sizes everywhere (a 2048x2048x512 volume has 2^31 voxels).
 *  - Every call returns a pk_status; on failure pk_last_error(ctx) holds a
 *    message.  Calls never abort the process.
 *  - Input buffers are borrowed read-only; outputs are caller-allocated.
 *  - Calls on one context are serialised by the context.
 *  - The GPU transforms support 2/3/5-smooth sizes on every axis (pad with
 *    pk_find_good_factor_number(n, 1, 0, 5) first); other sizes return
 *    PK_ERR_UNSUPPORTED_SIZE.
 */
#ifndef PKDECON_H
#define PKDECON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PK_OK = 0,
    PK_ERR_INVALID_ARG = 1,
    PK_ERR_OUT_SIZE = 2,          /* 'outSizeIsSmallerThanPsfSize' (decon_psf2otf.m:11-13) */
    PK_ERR_UNSUPPORTED_SIZE = 3,  /* an axis is not 2/3/5-smooth */
    PK_ERR_UNSUPPORTED_TYPE = 4,  /* input class other than uint16 / single */
    PK_ERR_HIP = 5,               /* HIP runtime error */
    PK_ERR_NO_MEMORY = 6,         /* device allocation failed */
    PK_ERR_BUDGET = 7,            /* request exceeds the context's HBM budget */
    PK_ERR_EMPTY_PSF = 8,         /* 'A valid nonzero PSF needs to be provided!' (omw_backprojector_generation.m:25-27) */
    PK_ERR_EMPTY_MASK = 9,        /* 'The OTF mask is empty ...' (omw_backprojector_generation.m:50-52) */
    PK_ERR_NO_DEVICE = 10,        /* no HIP device */
    PK_ERR_BUSY = 11,             /* serving queue full / job not finished */
    PK_ERR_CANCELLED = 12
} pk_status;

typedef enum { PK_UINT16 = 1, PK_SINGLE = 2, PK_UINT8 = 3 /* masks only */ } pk_dtype;
typedef enum { PK_MEM_HOST = 0, PK_MEM_DEVICE = 1 } pk_memspace;

/* A volume: data pointer (host or device), class, MATLAB size. */
typedef struct {
    void *data;
    int32_t dtype;   /* pk_dtype */
    int32_t mem;     /* pk_memspace */
    int64_t dims[3]; /* [n1 n2 n3], column-major */
} pk_vol;

/* Name-value options of decon_lucy_function / decon_lucy_omw_function
 * (decon_lucy_function.m:21-36; decon_lucy_omw_function.m:12-28). */
typedef struct {
    double background;     /* 'Background'  ([] -> 0) */
    double damp_factor;    /* 'dampFactor'  (1) */
    double scale_factor;   /* 'scaleFactor' (1) */
    double decon_offset;   /* 'deconOffset' (0) */
    int32_t edge_erosion;  /* 'EdgeErosion' (0) */
    int32_t save16bit;     /* 'save16bit'   (true): output class uint16, else single */
    int32_t has_bbox;      /* 'bbox' given */
    int32_t flags;         /* PK_FLAG_* (0 = the fused engine) */
    int64_t bbox[6];       /* 'bbox' [s1 s2 s3 e1 e2 e3], 1-based inclusive (decon_lucy_function.m:199-203) */
} pk_opts;

/* Comparison baselines (no counterpart in the reference's options):
 * PK_FLAG_UNFUSED runs the MATLAB-structured path -- one kernel per gpuArray
 * statement of decon_lucy_function.m:121-142 / decon_lucy_omw_function.m:105-112
 * (3-pass fftn / ifftn, separate OTF product, max, ./, .*) -- on the same
 * transforms; PK_FLAG_NO_OTF_CACHE rebuilds the OTFs on every call. */
#define PK_FLAG_UNFUSED 1
#define PK_FLAG_NO_OTF_CACHE 2
/* PK_FLAG_AUTOTUNE: measure the launch shapes and pass order of a padded size
 * the first time the context meets it (pk_autotune), then use them. */
#define PK_FLAG_AUTOTUNE 4
/* PK_FLAG_EPILOGUE_PASS: run the epilogue (damp, scale / offset / mask,
 * uint16, bbox) as a pass of its own after the last update instead of inside
 * it; the results are bitwise identical (comparison baseline). */
#define PK_FLAG_EPILOGUE_PASS 8
/* PK_FLAG_DEBUG: 'debug', true of decon_lucy_function / decon_lucy_omw_function
 * (decon_lucy_function.m:144-168, decon_lucy_omw_function.m:114-139): after
 * every iteration k the mean squared difference between the estimate and the
 * background-subtracted input is measured (on the transform-size volume), and
 * err_mat (NUMIT x 4, column-major) receives the reference's rows
 * [k, MSE_k, MSE_k / MSE_1, criterion] for k = 1..NUMIT, before the
 * reference's final trimming (RL: rows 1..istp, OMW: rows 1..floor(NUMIT /
 * saveStep)), which the caller applies.  Every saveStep-th estimate goes to
 * the iteration hook (pk_set_iter_hook; the reference writes it as
 * Iter_%04d.tif).  The iteration loop then runs kernel by kernel. */
#define PK_FLAG_DEBUG 16

typedef struct pk_ctx pk_ctx;

/* Defaults of the reference's inputParser. */
void pk_opts_default(pk_opts *o);

/* Context: one device, its stream, device workspace and the OTF cache (the
 * reference's `persistent H` / `persistent OTF_f OTF_b`,
 * decon_lucy_function.m:61-77, decon_lucy_omw_function.m:53-72).
 * hbm_budget_bytes = 0: no limit beyond the device. */
pk_status pk_ctx_create(int device, uint64_t hbm_budget_bytes, pk_ctx **out);
void pk_ctx_destroy(pk_ctx *ctx);
/* Drops cached OTFs and workspace (mirrors reset(gpuDevice),
 * demos/demo_RL_deconvolution.m:165-167). */
pk_status pk_reset(pk_ctx *ctx);
/* Use the caller's HIP stream (hipStream_t) for all work; NULL restores the
 * context's own stream. */
pk_status pk_set_stream(pk_ctx *ctx, void *hip_stream);
const char *pk_last_error(const pk_ctx *ctx);
/* Iteration hook of PK_FLAG_DEBUG calls (no counterpart in the reference,
 * which writes the estimates itself): called on the calling thread with the
 * estimate after iteration k (k a multiple of save_step), host single
 * values valid during the call, at the transform size dims.  save_step is
 * 'saveStep' (>= 1); hook may be NULL (metrics only). */
typedef void (*pk_iter_hook)(void *user, int32_t k, const float *estimate, const int64_t dims[3]);
pk_status pk_set_iter_hook(pk_ctx *ctx, int32_t save_step, pk_iter_hook hook, void *user);
/* Bytes of device memory the context holds (workspace + OTF cache). */
uint64_t pk_device_bytes(const pk_ctx *ctx);

/* [J, err_mat, k] = decon_lucy_function(I, PSF, NUMIT, ...)
 * (microscopeDataProcessing/decon/decon_lucy_function.m:1-215): Biggs-
 * accelerated Richardson-Lucy.  psf: host single, psf_dims.  out->dims must
 * equal the bbox size (or I's size); out->dtype = save16bit ? UINT16 : SINGLE.
 * err_mat (NUMIT x 4, column-major, may be NULL) receives zeros (debug off). */
pk_status pk_decon_lucy(pk_ctx *ctx, const pk_vol *I, const float *psf, const int64_t psf_dims[3], int32_t numit,
                        const pk_opts *opts, pk_vol *out, double *err_mat);

/* [J, err_mat] = decon_lucy_omw_function(I, PSF_f, PSF_b, NUMIT, ...)
 * (microscopeDataProcessing/decon/decon_lucy_omw_function.m:1-178).  The OTF
 * cache is keyed on both PSFs (the reference keys on PSF_b only). */
pk_status pk_decon_lucy_omw(pk_ctx *ctx, const pk_vol *I, const float *psf_f, const int64_t f_dims[3],
                            const float *psf_b, const int64_t b_dims[3], int32_t numit, const pk_opts *opts,
                            pk_vol *out, double *err_mat);

/* No counterpart in the reference: pk_decon_lucy_omw for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * pk_decon_lucy_omw, with these limits: I and out must be device memory
 * (PK_MEM_DEVICE); PK_FLAG_UNFUSED and PK_FLAG_AUTOTUNE are rejected; the
 * call allocates nothing and caches nothing -- its plan tables, OTFs,
 * estimate, spectrum and masks live in one static device array of
 * PK_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), and a size
 * needing more -- 8 bytes per voxel plus three half spectra of
 * 8 * roundup(n1/2 + 1, 16) * n2 * n3 bytes, 2 bytes more per voxel with
 * EdgeErosion (about 20 and 22 bytes per voxel; 192x192x160 fits) -- returns
 * PK_ERR_UNSUPPORTED_SIZE.  Calls that overlap on one GPU share that array
 * and race: serialise them. */
#define PK_NOALLOC_SCRATCH_BYTES (128ull << 20)
pk_status pk_decon_lucy_omw_noalloc(pk_ctx *ctx, const pk_vol *I, const float *psf_f, const int64_t f_dims[3],
                                    const float *psf_b, const int64_t b_dims[3], int32_t numit, const pk_opts *opts,
                                    pk_vol *out, double *err_mat);

/* The whole-volume edge-erosion mask of the large-volume driver
 * (microscopeDataProcessing/decon/RLdecon_large_in_memory.m:150-155):
 * bw = im > 0 with the six faces cleared, mask = bwdist(~bw) > EdgeErosion - 1
 * (Euclidean distance in voxels to the nearest false voxel of the volume).
 * im: uint16 or single, host or device; mask: PK_UINT8 (0 / 1), same dims,
 * host or device.  1 <= edge_erosion <= 256. */
pk_status pk_edge_erosion_mask(pk_ctx *ctx, const pk_vol *im, int32_t edge_erosion, pk_vol *mask);

/* One chunk of a batch: input and output volumes of the same (any) size. */
typedef struct {
    pk_vol in;
    pk_vol out;
} pk_batch_item;

/* Batched deconvolution of n chunks of different sizes sharing the PSFs and
 * options: the per-chunk loop of RLdecon_large_in_memory.m:190-260 /
 * RLdecon_for_zarr_block.m:101-171 around decon_lucy_function (method 0) or
 * decon_lucy_omw_function (method 1).  Chunks whose size is not 2/3/5-smooth
 * are zero-padded to the next smooth size and cropped back; a PSF larger than
 * a chunk is centre-cropped to it first (RLdecon.m:264-270, 326-332).  Every
 * pass runs as one launch over all chunks; each chunk's result is bitwise the
 * same as a single pk_decon_lucy / pk_decon_lucy_omw call on the padded chunk.
 * bbox and EdgeErosion are per-chunk options and are rejected here. */
pk_status pk_decon_batch(pk_ctx *ctx, int32_t method, const pk_batch_item *items, int32_t n, const float *psf_f,
                         const int64_t f_dims[3], const float *psf_b, const int64_t b_dims[3], int32_t numit,
                         const pk_opts *opts);

/* ---- Serving: a request loop on one device (no counterpart in the reference;
 * its nearest analogue is the per-file loop of XR_decon_data_wrapper.m:332-511).
 * Continuous batching: requests are admitted FIFO as memory under the HBM
 * budget given to pk_ctx_create allows, every admitted request advances one
 * iteration per step in batched launches shared with the others, and a
 * request's result is identical alone and in any mix.  A request larger than
 * the budget fails with PK_ERR_BUDGET once nothing else is running.  The
 * caller's buffers must stay valid until the request is DONE, FAILED or
 * CANCELLED.  bbox and EdgeErosion are not accepted for served requests. */
typedef struct {
    int32_t method;          /* 0 = decon_lucy_function (Biggs RL), 1 = decon_lucy_omw_function */
    int32_t numit;
    pk_vol in;
    pk_vol out;              /* same size as in; class per opts.save16bit */
    const float *psf_f;      /* copied at submit */
    int64_t f_dims[3];
    const float *psf_b;      /* method 1 only; copied at submit */
    int64_t b_dims[3];
    pk_opts opts;
} pk_request;

typedef enum { PK_REQ_QUEUED = 0, PK_REQ_RUNNING = 1, PK_REQ_DONE = 2, PK_REQ_FAILED = 3, PK_REQ_CANCELLED = 4 } pk_req_state;

typedef struct {
    int32_t state;            /* pk_req_state */
    int32_t status;           /* pk_status of a failed request */
    int32_t iterations_done;
    int32_t reserved;
    double latency_ms;        /* submit -> output ready (or -> now while pending) */
    double queue_ms;          /* submit -> admission */
} pk_req_info;

typedef struct {
    uint64_t submitted, completed, failed;
    int32_t active, queued;
    double voxels;            /* voxels of completed requests */
} pk_serve_stats;

/* max_active: requests in flight at most (<= 0: 32); the rest wait in FIFO order. */
pk_status pk_serve_start(pk_ctx *ctx, int32_t max_active);
/* Stops after the requests in flight finish; queued requests are cancelled. */
pk_status pk_serve_stop(pk_ctx *ctx);
pk_status pk_serve_submit(pk_ctx *ctx, const pk_request *req, uint64_t *id);
/* wait_ms: 0 = poll, < 0 = wait until finished, > 0 = wait at most that long. */
pk_status pk_serve_wait(pk_ctx *ctx, uint64_t id, double wait_ms, pk_req_info *info);
/* Cancels a queued request (PK_ERR_BUSY once it runs). */
pk_status pk_serve_cancel(pk_ctx *ctx, uint64_t id);
/* Forgets a finished request. */
pk_status pk_serve_release(pk_ctx *ctx, uint64_t id);
pk_status pk_serve_stats_get(pk_ctx *ctx, pk_serve_stats *out);

/* otf = decon_psf2otf(psf, outSize) (microscopeDataProcessing/decon/decon_psf2otf.m:1-46).
 * otf: host buffer of prod(out_dims) complex single values, interleaved
 * (re, im), the full spectrum as MATLAB returns it. */
pk_status pk_psf2otf(pk_ctx *ctx, const float *psf, const int64_t psf_dims[3], const int64_t out_dims[3],
                     float *otf);

/* Name-value options of omw_backprojector_generation
 * (microscopeDataProcessing/decon/omw_backprojector_generation.m:7-15). */
typedef struct {
    double otf_cum_thresh;   /* 'OTFCumThresh'  (0.9) */
    double otf_area_thresh;  /* 'OTFAreaThresh' (100) */
    double hann_win_lo;      /* 'hannWinBounds' (1) (0.8) */
    double hann_win_hi;      /* 'hannWinBounds' (2) (1.0) */
    double min_int_thresh;   /* 'minIntThrsh'   (1e-3) */
} pk_bp_opts;
void pk_bp_opts_default(pk_bp_opts *o);

/* [b_omw, OTF_bp_omw, abs_OTF_c, OTF_mask] = omw_backprojector_generation(psf, alpha, skewed, ...)
 * (microscopeDataProcessing/decon/omw_backprojector_generation.m:1-166).
 * psf: host single, psf_dims; skewed: -1 = [] (auto-detect), 0 = false,
 * 1 = true.  Outputs are host buffers of prod(psf_dims) elements: b_omw
 * (single), otf_bp_omw (complex single, interleaved; may be NULL),
 * abs_otf_c (single; may be NULL), otf_mask (single 0/1, or 0..3 counts for
 * skewed data; may be NULL).  skewed_used receives the branch taken.  The
 * nearest-edge search runs on the context's GPU. */
pk_status pk_omw_backprojector(pk_ctx *ctx, const float *psf, const int64_t psf_dims[3], double alpha,
                               int32_t skewed, const pk_bp_opts *opts, float *b_omw, float *otf_bp_omw,
                               float *abs_otf_c, float *otf_mask, int32_t *skewed_used);

/* [PSF_bp, OTF_bp] = BackProjector(PSF_fp, bp_type, alpha, beta, n, resFlag, iRes)
 * (third_parties/WBDeconvolution/BackProjector.m:1-233): traditional (flipped
 * PSF), gaussian, butterworth, wiener and wiener-butterworth back projectors.
 * alpha == 1 / beta == 1 take the traditional back projector's cutoff gain, as
 * the reference does; res_flag 0/1/2 = FWHM/sqrt(2) / FWHM / ires (pixels).
 * Host computation (no device use; ctx may be NULL, then no error text).
 * psf_fp: host single, dims; psf_bp: host double, prod(dims); otf_bp: host
 * complex double interleaved, 2*prod(dims) (may be NULL); info (may be NULL):
 * 10 doubles {FWHMx, FWHMy, FWHMz, tx, ty, tz, beta_fp, alpha, beta, beta_wienerx}. */
enum { PK_BP_TRADITIONAL = 0, PK_BP_GAUSSIAN = 1, PK_BP_BUTTERWORTH = 2, PK_BP_WIENER = 3, PK_BP_WIENER_BUTTERWORTH = 4 };
pk_status pk_wb_backprojector(pk_ctx *ctx, const float *psf_fp, const int64_t dims[3], int32_t bp_type, double alpha,
                              double beta, int32_t n, int32_t res_flag, const double ires[3], double *psf_bp,
                              double *otf_bp, double *info);

/* Launch-shape autotuner (no counterpart in the reference, whose FFTs are
 * MATLAB's): times the candidate row-pass values-per-thread counts, the
 * strided passes' values per thread, block sizes and tile widths, and the pass
 * order (which strided axis carries the fused OTF convolution) for the padded
 * size dims on the context's GPU, and keeps the fastest for later calls of that
 * size.  Results are shared by the process's contexts and appended to the file
 * named by the environment variable PK_TUNE_FILE when set (read back at the
 * next plan of the size).  report (may be NULL) receives a one-line summary.
 * Call it before serving: requests in flight keep the shapes they started with
 * only when nothing retunes their size. */
pk_status pk_autotune(pk_ctx *ctx, const int64_t dims[3], char *report, size_t report_len);

/* Sets the launch shapes of a padded size without measuring (no counterpart
 * in the reference): shapes = {row values per thread, dim-2 values per
 * thread, block size, tile width, dim-3 values per thread, block size, tile
 * width, pass order (2 or 3: the strided axis carrying the OTF convolution)},
 * the fields of one PK_TUNE_FILE line.  Applies to this context's later calls
 * and to plans the process's other contexts make afterwards; not written to
 * the file.  PK_ERR_INVALID_ARG when a shape is not feasible for the size. */
pk_status pk_tune_set(pk_ctx *ctx, const int64_t dims[3], const int32_t shapes[8]);

/* Transform diagnostics (MATLAB fftn / ifftn of real data, decon_lucy_function.m:136,142).
 * Device pointers.  r2c: real N floats -> half spectrum (floor(n1/2)+1)*n2*n3
 * complex (interleaved); c2r: half spectrum -> real, unnormalised (N * ifftn). */
pk_status pk_fft_r2c(pk_ctx *ctx, const float *d_real, float *d_spec, const int64_t dims[3]);
pk_status pk_fft_c2r(pk_ctx *ctx, const float *d_spec, float *d_real, const int64_t dims[3]);

/* Per-pass timing (no counterpart in the reference, which only has tic/toc):
 * HIP events recorded around every kernel launch on the context's stream.
 * enable=1 clears the table and starts recording, 0 stops.  Entries are
 * aggregated by pass name with the pass's algorithmic bytes per launch
 * (every operand read or written once). */
typedef struct {
    char name[32];
    int64_t launches;
    double total_ms;
    double bytes_per_launch;
} pk_prof_entry;
pk_status pk_profile_enable(pk_ctx *ctx, int32_t enable);
/* Copies up to max entries into out; returns the number of entries (-1: null ctx). */
int32_t pk_profile_read(pk_ctx *ctx, pk_prof_entry *out, int32_t max);

/* Copy probes (diagnostics; no counterpart in the reference): copies the
 * pitched half spectrum of a dims-sized transform into a second device
 * buffer with one pass's access pattern and no arithmetic -- pattern 0:
 * sequential 16-byte accesses; 1: the dim-2 column passes (tiles of `tile`
 * adjacent columns, column stride = row pitch); 2: the dim-3 column pass
 * (column stride = plane pitch); 3: pattern 2 with a second input read at the
 * same points (the OTF convolution's traffic).  Each thread moves vals_per_thread points
 * (4, 8, 16 or 32; the column length must be a multiple) of one column.
 * Column patterns store with the strided passes' cache policy (nontemporal
 * stores for tiles of 16 or more columns by default); the sequential copy is
 * a plain one.  flags: bit 0 takes
 * the passes' XCD-aware block order; bit 1 runs each copy cold -- a 1 GiB
 * device buffer (four times the memory-side cache) is read before it,
 * outside the timed interval -- instead of back to back, where a copy can
 * read what the previous one left in that cache.  *ms = mean time of reps
 * copies (HIP events on the context's stream), *bytes = bytes read plus
 * written per copy (valid columns only).  Uses the context's workspace. */
pk_status pk_probe_copy(pk_ctx *ctx, int32_t pattern, const int64_t dims[3], int32_t tile, int32_t vals_per_thread,
                        int32_t flags, int32_t reps, double *ms, double *bytes);

/* findGoodFactorNumber(givenNum, directionFlag, allowOddSeven) (utils/findGoodFactorNumber.m:1-26).
 * max_prime = 7 is the reference rule; 5 gives the 2/3/5-smooth sizes the GPU transforms need. */
int64_t pk_find_good_factor_number(int64_t n, int32_t direction, int32_t allow_odd_seven, int32_t max_prime);

/* deskewFrame3D(vol, angle, dz, xyPixelSize, reverse, 'Crop', crop) with the
 * default linear interpolation (pointDetection/llsmtools/transforms/
 * deskewFrame3D.m:26-70, applied to the deconvolved volume at
 * microscopeDataProcessing/decon/RLdecon.m:448-449).  pk_deskew_out_dims gives
 * the output size [ny, nxOut, nz] (host only); pk_deskew_frame3d writes it
 * from the single-precision volume at d_vol into d_out (both device memory,
 * column-major) and returns when done. */
pk_status pk_deskew_out_dims(const int64_t dims[3], double angle, double dz, double xy_pixel_size, int32_t reverse,
                             int32_t crop, int64_t out_dims[3]);
pk_status pk_deskew_frame3d(pk_ctx *ctx, const float *d_vol, const int64_t dims[3], double angle, double dz,
                            double xy_pixel_size, int32_t reverse, int32_t crop, float *d_out);

/* volout = deskewRotateFrame3D(vol, angle, dz, xyPixelSize, 'reverse', reverse,
 * 'objectiveScan', objective_scan, 'xStepThresh', x_step_thresh, 'bbox', bbox)
 * with the default linear interpolation and no resampling
 * (microscopeDataProcessing/deskew_rotate/deskewRotateFrame3D.m:50-253, the
 * combined deskew and rotation RLdecon.m:489-495 runs by default).  When the
 * slice-to-slice shift cos(angle) dz / xyPixelSize exceeds x_step_thresh
 * (2.0 by default) the slices are first interpolated in skewed space to a
 * finer step (skewed_space_interp_defined_stepsize.m:24-100); a stage scan is
 * then warped row by row as volume_deskew_rotate_warp_mex.cpp:217-282 does, an
 * objective scan through imwarp's trilinear form.  in / out: device volumes;
 * in uint16 or single; out single, or uint16 = uint16() of the single result;
 * out->dims = the bbox size ([s1 s2 s3 e1 e2 e3], 1-based inclusive; NULL =
 * the whole output, pk_deskew_rotate_out_dims).  Returns when done. */
pk_status pk_deskew_rotate_out_dims(const int64_t dims[3], double angle, double dz, double xy_pixel_size,
                                    int32_t reverse, int32_t objective_scan, double x_step_thresh, int64_t out_dims[3]);
pk_status pk_deskew_rotate_frame3d(pk_ctx *ctx, const pk_vol *in, double angle, double dz, double xy_pixel_size,
                                   int32_t reverse, int32_t objective_scan, double x_step_thresh, const int64_t bbox[6],
                                   pk_vol *out);
/* The same with interpMethod 'cubic' (deskewRotateFrame3D.m:73 leaves the fast
 * path, :98-110 and :209-210): for a stage scan whose step exceeds
 * x_step_thresh the skewed-space interpolation runs first (linear; its result
 * rounded to uint16 values when in and out are uint16, as the reference's
 * interpolation returns uint16 with save16bit), then imwarp(vol_1,
 * affine3d(ds_S*(T1*S*R*T2)), 'cubic', 'FillValues', 0, 'OutputView', RA) as
 * pk_affine_warp3d_cubic computes it; bbox is RA's output window (:203-205).
 * imwarp returns the class of vol_1, so a caller mirroring the reference
 * passes a uint16 out for a uint16 volume that is not interpolated, or that
 * is interpolated with save16bit. */
pk_status pk_deskew_rotate_frame3d_cubic(pk_ctx *ctx, const pk_vol *in, double angle, double dz, double xy_pixel_size,
                                         int32_t reverse, int32_t objective_scan, double x_step_thresh,
                                         const int64_t bbox[6], pk_vol *out);
/* deskewRotateFrame3D(vol, angle, dz, xyPixelSize, ..., 'resampleFactor',
 * resample, 'interpMethod', linear (interp 0) or cubic (1))
 * (deskewRotateFrame3D.m:72-73, 183-210): a resampleFactor other than
 * [1 1 1] (x y z order) leaves the fast path; the skewed-space interpolation
 * runs first when the step exceeds x_step_thresh (its result rounded to
 * uint16 values when in and out are uint16), then imwarp(vol_1,
 * affine3d(ds_S*(T1*S*R*T2)*(RT1*RS*RT2)), interpMethod, 'FillValues', 0,
 * 'OutputView', RA) with RS = diag(1 ./ resample) about the box centre and RA
 * of size round(outSize ./ resample([2 1 3])) (or the bbox window of it).
 * resample NULL or [1 1 1]: pk_deskew_rotate_frame3d (interp 0) or
 * pk_deskew_rotate_frame3d_cubic (interp 1), unless interp + 2 is passed:
 * that takes the imwarp form all the same, as the reference does for
 * 'gpuProcess' true (:73).  imwarp returns the class of
 * vol_1: a uint16 out for a uint16 volume that is not interpolated, or that is
 * interpolated with save16bit.  pk_deskew_rotate_resample_out_dims gives the
 * output size (host only). */
pk_status pk_deskew_rotate_resample_out_dims(const int64_t dims[3], double angle, double dz, double xy_pixel_size,
                                             int32_t reverse, int32_t objective_scan, double x_step_thresh,
                                             const double resample[3], int64_t out_dims[3]);
pk_status pk_deskew_rotate_frame3d_resample(pk_ctx *ctx, const pk_vol *in, double angle, double dz,
                                            double xy_pixel_size, int32_t reverse, int32_t objective_scan,
                                            double x_step_thresh, const double resample[3], int32_t interp,
                                            const int64_t bbox[6], pk_vol *out);

/* Affine resampling as imwarp(vol, affine3d(T), 'linear', 'FillValues', 0,
 * 'OutputView', imref3d(out_dims, 1, 1, 1)) does for rotateFrame3D
 * (pointDetection/llsmtools/transforms/rotateFrame3D.m:46-100, used after
 * deskew at microscopeDataProcessing/decon/RLdecon.m:483-485): inv_tform is
 * inv(T), row-major 4x4 in the row-vector convention ([x y z 1] = [u v w 1] *
 * inv_tform, x = column, y = row, z = page, 1-based); trilinear over the
 * input padded with one voxel of zeros.  d_vol / d_out are device memory,
 * column-major single; returns when done. */
pk_status pk_affine_warp3d(pk_ctx *ctx, const float *d_vol, const int64_t dims[3], const double inv_tform[16],
                           const int64_t out_dims[3], float *d_out);
/* The same resampling with imwarp's 'cubic' method (deskewFrame3D.m:70,
 * rotateFrame3D.m:96 with 'Interp', 'cubic'): the input padded with one voxel
 * of zeros, cubic convolution (Keys, a = -0.5) over the padded grid with Keys'
 * boundary rule f(-1) = 3 f(0) - 3 f(1) + f(2) for the taps past its ends,
 * fill value 0 outside the padded grid.  Every input dimension >= 2. */
pk_status pk_affine_warp3d_cubic(pk_ctx *ctx, const float *d_vol, const int64_t dims[3], const double inv_tform[16],
                                 const int64_t out_dims[3], float *d_out);

/* TIFF LZW strip codec (host code; no context, no device).  Decodes one LZW
 * strip as libtiff's codec does for the reference reader's
 * TIFFReadEncodedStrip calls (io/cpp-tiff/src/parallelreadtiff.cpp:283-300),
 * and encodes one strip as the reference writer's encoder
 * (io/cpp-tiff/src/lzwencode.cpp, used per strip by parallelwritetiff.cpp:
 * 52-60).  *out_len receives the bytes written; PK_ERR_OUT_SIZE when cap is too
 * small, PK_ERR_INVALID_ARG on a malformed strip.  Thread-safe: strips are
 * independent.  pk_lzw_encode_bound(n) is a capacity that always suffices. */
pk_status pk_lzw_decode(const uint8_t *src, size_t n, uint8_t *dst, size_t cap, size_t *out_len);
pk_status pk_lzw_encode(const uint8_t *src, size_t n, uint8_t *dst, size_t cap, size_t *out_len);
size_t pk_lzw_encode_bound(size_t n);

/* LZ4 block codec (host code; no context, no device): the streams inside the
 * blosc-compressed Zarr chunks the reference reads and writes by default
 * (blosc + lz4, io/cpp-zarr/src/zarr.cpp:15-18, through c-blosc in
 * parallelreadzarr.cpp:148-157).  Same return conventions as the LZW codec;
 * pk_lz4_encode needs cap >= pk_lz4_encode_bound(n). */
pk_status pk_lz4_decode(const uint8_t *src, size_t n, uint8_t *dst, size_t cap, size_t *out_len);
pk_status pk_lz4_encode(const uint8_t *src, size_t n, uint8_t *dst, size_t cap, size_t *out_len);
size_t pk_lz4_encode_bound(size_t n);

#ifdef __cplusplus
}
#endif
#endif /* PKDECON_H */
