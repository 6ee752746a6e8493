/*
 * pk_oracle.h -- CPU restatement of PetaKit5D's RL / OMW deconvolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in petakit5d_amd/: only tests/, the smoke() entry and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Arrays follow MATLAB conventions: column-major, dim 1 contiguous, complex
 * data interleaved (re, im).  Every function has an _f32 variant (single
 * precision arithmetic, the reference's CPU path) and an _f64 variant (the
 * FP64 oracle).  Function bodies live in pk_oracle_impl.h.
 *
 * Reference files restated (paths relative to the PetaKit5D tree):
 *   microscopeDataProcessing/decon/decon_psf2otf.m
 *   microscopeDataProcessing/decon/decon_lucy_function.m
 *   microscopeDataProcessing/decon/decon_lucy_omw_function.m
 *   microscopeDataProcessing/decon/decon_mask_edge_erosion.m
 *   utils/findGoodFactorNumber.m
 */
#ifndef PK_ORACLE_H
#define PK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PKO_OK 0
#define PKO_ERR_ARG 1
#define PKO_ERR_OUTSIZE 2   /* 'outSizeIsSmallerThanPsfSize' (decon_psf2otf.m:12) */
#define PKO_ERR_NOMEM 3

typedef struct {
    double background;   /* 'Background'  (decon_lucy_function.m:26) */
    double dampFactor;   /* 'dampFactor'  (:29) */
    double scaleFactor;  /* 'scaleFactor' (:30) */
    double deconOffset;  /* 'deconOffset' (:31) */
    int32_t edgeErosion; /* 'EdgeErosion' (:32) */
    int32_t save16bit;   /* 'save16bit'   (:28) */
} pko_opts;

/* In-place 3-D complex FFT.  inverse=0: unnormalised forward (fftn);
 * inverse=1: backward scaled by 1/N (ifftn).  Any size (radix 2/3/4/5 plus a
 * direct DFT for other primes).  nthreads<=0 uses the OpenMP default. */
int pko_fftn_f32(float *data, int64_t n1, int64_t n2, int64_t n3, int inverse);
int pko_fftn_f64(double *data, int64_t n1, int64_t n2, int64_t n3, int inverse);

/* decon_psf2otf(psf, outSize): full complex OTF (N complex, interleaved). */
int pko_psf2otf_f32(const float *psf, const int64_t pdims[3], const int64_t odims[3], float *otf);
int pko_psf2otf_f64(const double *psf, const int64_t pdims[3], const int64_t odims[3], double *otf);

/* decon_lucy_function / decon_lucy_omw_function, CPU branch (useGPU=false).
 * I: single(I) (or double) at the transform size.  J_out: the estimate after
 * the epilogue's damp / scale / offset / mask steps but before uint16
 * conversion (u16_out, if non-NULL, receives uint16(J) when save16bit).
 * I_raw_nonzero: optional mask I~=0 of the raw input (NULL = derive from I). */
int pko_decon_lucy_f32(const float *I, const int64_t dims[3], const float *psf, const int64_t pdims[3],
                       int numit, const pko_opts *o, float *J_out, uint16_t *u16_out);
int pko_decon_lucy_f64(const double *I, const int64_t dims[3], const double *psf, const int64_t pdims[3],
                       int numit, const pko_opts *o, double *J_out, uint16_t *u16_out);
int pko_decon_omw_f32(const float *I, const int64_t dims[3], const float *psf_f, const int64_t pfdims[3],
                      const float *psf_b, const int64_t pbdims[3], int numit, const pko_opts *o,
                      float *J_out, uint16_t *u16_out);
int pko_decon_omw_f64(const double *I, const int64_t dims[3], const double *psf_f, const int64_t pfdims[3],
                      const double *psf_b, const int64_t pbdims[3], int numit, const pko_opts *o,
                      double *J_out, uint16_t *u16_out);

/* decon_mask_edge_erosion(mask, E) (decon_mask_edge_erosion.m:1-30) */
int pko_mask_edge_erosion(uint8_t *mask, const int64_t dims[3], int E);

/* findGoodFactorNumber(n, dir, allowOddSeven) (findGoodFactorNumber.m:1-26);
 * max_prime=7 is the reference rule, max_prime=5 the 2/3/5-smooth variant. */
int64_t pko_find_good_factor_number(int64_t n, int dir, int allow_odd_seven, int max_prime);

/* MATLAB uint16(x): round half away from zero, saturate, NaN -> 0. */
uint16_t pko_to_u16(double x);

/* Neighbourhood of strel('disk', r) (default N = 4): half-width w[dy + R] of
 * every row dy in [-R, R]; returns R (w holds 2r + 1 ints), -1 for r < 0. */
int pko_strel_disk(int r, int *w);

#ifdef __cplusplus
}
#endif
#endif
