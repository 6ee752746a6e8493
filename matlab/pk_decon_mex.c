/*
 * pk_decon_mex.c -- MATLAB MEX gateway over libpkdecon.so (include/pkdecon.h).
 *
 *   J = pk_decon_mex('rl',  I, PSF, NUMIT, opts)          % decon_lucy_function
 *   J = pk_decon_mex('omw', I, PSF_f, PSF_b, NUMIT, opts) % decon_lucy_omw_function
 *   [b, otf_bp, abs_otf_c, mask] = pk_decon_mex('bp', psf, alpha, skewed, bpopts)
 *   [psf_bp, otf_bp] = pk_decon_mex('wbbp', psf, bp_type, alpha, beta, n, resFlag, iRes)  % BackProjector
 *   otf = pk_decon_mex('psf2otf', psf, outSize)
 *   n = pk_decon_mex('fgfn', n, dir, allowOddSeven, maxPrime)
 *   pk_decon_mex('reset')
 *
 * Conventions follow microscopeDataProcessing/crop/crop3d_mex.c:33-113:
 * arity checks, mexErrMsgIdAndTxt identifiers, dispatch on mxGetClassID,
 * outputs created with mxCreateNumericArray.  Inputs are borrowed read-only
 * (unlike indexing3d_mex.c:20, nothing is written in place).  The device
 * context is created on first use and destroyed at mexAtExit.
 *
 * Build (MATLAB):  mex -R2018a pk_decon_mex.c -I../include -L../petakit5d_amd -lpkdecon
 */
#include <stdint.h>
#include <string.h>

#include "mex.h"
#include "pkdecon.h"

static pk_ctx *g_ctx = NULL;

static void pk_mex_cleanup(void)
{
    if (g_ctx) pk_ctx_destroy(g_ctx);
    g_ctx = NULL;
}

static pk_ctx *ctx_get(void)
{
    if (!g_ctx) {
        pk_status st = pk_ctx_create(0, 0, &g_ctx);
        if (st != PK_OK) mexErrMsgIdAndTxt("pkdecon:noDevice", "pk_ctx_create failed (status %d)", (int)st);
        mexAtExit(pk_mex_cleanup);
    }
    return g_ctx;
}

static void check(pk_status st)
{
    static const char *ids[] = {"ok", "invalidArg", "outSizeIsSmallerThanPsfSize", "unsupportedSize",
                                "unsupportedType", "hip", "noMemory", "budget", "emptyPsf", "emptyMask",
                                "noDevice", "busy", "cancelled"};
    if (st == PK_OK) return;
    char id[64];
    strcpy(id, "pkdecon:");
    strncat(id, (st >= 0 && st <= 12) ? ids[st] : "error", sizeof(id) - 10);
    mexErrMsgIdAndTxt(id, "%s", pk_last_error(g_ctx));
}

static void dims3(const mxArray *a, int64_t d[3])
{
    mwSize nd = mxGetNumberOfDimensions(a);
    const mwSize *dd = mxGetDimensions(a);
    if (nd > 3) mexErrMsgIdAndTxt("pkdecon:invalidArg", "at most 3-D arrays");
    for (int i = 0; i < 3; ++i) d[i] = i < (int)nd ? (int64_t)dd[i] : 1;
}

static double field_or(const mxArray *s, const char *name, double def)
{
    if (!s || !mxIsStruct(s)) return def;
    const mxArray *f = mxGetField(s, 0, name);
    if (!f || mxIsEmpty(f)) return def;
    return mxGetScalar(f);
}

static void read_opts(const mxArray *s, pk_opts *o)
{
    pk_opts_default(o);
    o->background = field_or(s, "Background", 0);
    o->damp_factor = field_or(s, "dampFactor", 1);
    o->scale_factor = field_or(s, "scaleFactor", 1);
    o->decon_offset = field_or(s, "deconOffset", 0);
    o->edge_erosion = (int32_t)field_or(s, "EdgeErosion", 0);
    o->save16bit = field_or(s, "save16bit", 1) != 0;
    if (s && mxIsStruct(s)) {
        const mxArray *b = mxGetField(s, 0, "bbox");
        if (b && !mxIsEmpty(b)) {
            if (mxGetNumberOfElements(b) != 6) mexErrMsgIdAndTxt("pkdecon:invalidArg", "bbox must have 6 elements");
            o->has_bbox = 1;
            for (int i = 0; i < 6; ++i) o->bbox[i] = (int64_t)((const double *)mxGetPr(b))[i];
        }
    }
}

static const float *single_arg(const mxArray *a, const char *what)
{
    if (!mxIsSingle(a)) mexErrMsgIdAndTxt("pkdecon:invalidArg", "%s must be single", what);
    return (const float *)mxGetData(a);
}

static void decon(int method, int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const int need = method == 1 ? 6 : 5;
    if (nrhs != need) mexErrMsgIdAndTxt("pkdecon:invalidArg", "wrong number of inputs");
    const mxArray *I = prhs[1];
    pk_vol vin;
    memset(&vin, 0, sizeof(vin));
    switch (mxGetClassID(I)) {
        case mxUINT16_CLASS: vin.dtype = PK_UINT16; break;
        case mxSINGLE_CLASS: vin.dtype = PK_SINGLE; break;
        default: mexErrMsgIdAndTxt("pkdecon:unsupportedType", "I must be uint16 or single");
    }
    vin.data = mxGetData(I);
    vin.mem = PK_MEM_HOST;
    dims3(I, vin.dims);
    int64_t fd[3], bd[3];
    const float *pf = single_arg(prhs[2], "PSF");
    dims3(prhs[2], fd);
    const float *pb = NULL;
    if (method == 1) {
        pb = single_arg(prhs[3], "PSF_b");
        dims3(prhs[3], bd);
    }
    const int32_t numit = (int32_t)mxGetScalar(prhs[method == 1 ? 4 : 3]);
    pk_opts o;
    read_opts(prhs[need - 1], &o);
    mwSize od[3] = {(mwSize)vin.dims[0], (mwSize)vin.dims[1], (mwSize)vin.dims[2]};
    if (o.has_bbox)
        for (int i = 0; i < 3; ++i) od[i] = (mwSize)(o.bbox[i + 3] - o.bbox[i] + 1);
    plhs[0] = mxCreateNumericArray(3, od, o.save16bit ? mxUINT16_CLASS : mxSINGLE_CLASS, mxREAL);
    pk_vol vout;
    vout.data = mxGetData(plhs[0]);
    vout.dtype = o.save16bit ? PK_UINT16 : PK_SINGLE;
    vout.mem = PK_MEM_HOST;
    for (int i = 0; i < 3; ++i) vout.dims[i] = (int64_t)od[i];
    double *err = NULL;
    if (nlhs > 1) {
        plhs[1] = mxCreateDoubleMatrix(numit > 0 ? numit : 0, 4, mxREAL);
        err = mxGetPr(plhs[1]);
    }
    pk_ctx *c = ctx_get();
    check(method == 1 ? pk_decon_lucy_omw(c, &vin, pf, fd, pb, bd, numit, &o, &vout, err)
                      : pk_decon_lucy(c, &vin, pf, fd, numit, &o, &vout, err));
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgIdAndTxt("pkdecon:invalidArg", "first input is the operation name");
    char *op = mxArrayToString(prhs[0]);
    if (!strcmp(op, "rl")) {
        decon(0, nlhs, plhs, nrhs, prhs);
    } else if (!strcmp(op, "omw")) {
        decon(1, nlhs, plhs, nrhs, prhs);
    } else if (!strcmp(op, "fgfn")) {
        if (nrhs != 5) mexErrMsgIdAndTxt("pkdecon:invalidArg", "fgfn takes n, dir, allowOddSeven, maxPrime");
        const int64_t r = pk_find_good_factor_number((int64_t)mxGetScalar(prhs[1]), (int32_t)mxGetScalar(prhs[2]),
                                                     (int32_t)mxGetScalar(prhs[3]), (int32_t)mxGetScalar(prhs[4]));
        plhs[0] = mxCreateDoubleScalar((double)r);
    } else if (!strcmp(op, "reset")) {
        if (g_ctx) check(pk_reset(g_ctx));
    } else if (!strcmp(op, "psf2otf")) {
        if (nrhs != 3) mexErrMsgIdAndTxt("pkdecon:invalidArg", "psf2otf takes psf, outSize");
        int64_t pd[3], od[3];
        dims3(prhs[1], pd);
        const double *os = mxGetPr(prhs[2]);
        const size_t no = mxGetNumberOfElements(prhs[2]);
        for (int i = 0; i < 3; ++i) od[i] = i < (int)no ? (int64_t)os[i] : 1;
        mwSize odm[3] = {(mwSize)od[0], (mwSize)od[1], (mwSize)od[2]};
        plhs[0] = mxCreateNumericArray(3, odm, mxSINGLE_CLASS, mxCOMPLEX);
        float *tmp = (float *)mxMalloc(sizeof(float) * 2 * (size_t)(od[0] * od[1] * od[2]));
        check(pk_psf2otf(ctx_get(), single_arg(prhs[1], "psf"), pd, od, tmp));
        memcpy(mxGetData(plhs[0]), tmp, sizeof(float) * 2 * (size_t)(od[0] * od[1] * od[2]));  /* interleaved (R2018a) */
        mxFree(tmp);
    } else if (!strcmp(op, "bp")) {
        if (nrhs != 5) mexErrMsgIdAndTxt("pkdecon:invalidArg", "bp takes psf, alpha, skewed, bpopts");
        int64_t pd[3];
        dims3(prhs[1], pd);
        pk_bp_opts bo;
        pk_bp_opts_default(&bo);
        bo.otf_cum_thresh = field_or(prhs[4], "OTFCumThresh", bo.otf_cum_thresh);
        bo.otf_area_thresh = field_or(prhs[4], "OTFAreaThresh", bo.otf_area_thresh);
        bo.min_int_thresh = field_or(prhs[4], "minIntThrsh", bo.min_int_thresh);
        const mxArray *hw = mxIsStruct(prhs[4]) ? mxGetField(prhs[4], 0, "hannWinBounds") : NULL;
        if (hw && mxGetNumberOfElements(hw) == 2) {
            bo.hann_win_lo = mxGetPr(hw)[0];
            bo.hann_win_hi = mxGetPr(hw)[1];
        }
        const int32_t sk = mxIsEmpty(prhs[3]) ? -1 : (mxGetScalar(prhs[3]) != 0);
        mwSize dm[3] = {(mwSize)pd[0], (mwSize)pd[1], (mwSize)pd[2]};
        plhs[0] = mxCreateNumericArray(3, dm, mxSINGLE_CLASS, mxREAL);
        mxArray *otf = mxCreateNumericArray(3, dm, mxSINGLE_CLASS, mxCOMPLEX);
        mxArray *absc = mxCreateNumericArray(3, dm, mxSINGLE_CLASS, mxREAL);
        mxArray *mask = mxCreateNumericArray(3, dm, mxSINGLE_CLASS, mxREAL);
        int32_t used = 0;
        check(pk_omw_backprojector(ctx_get(), single_arg(prhs[1], "psf"), pd, mxGetScalar(prhs[2]), sk, &bo,
                                   (float *)mxGetData(plhs[0]), (float *)mxGetData(otf), (float *)mxGetData(absc),
                                   (float *)mxGetData(mask), &used));
        if (nlhs > 1) plhs[1] = otf; else mxDestroyArray(otf);
        if (nlhs > 2) plhs[2] = absc; else mxDestroyArray(absc);
        if (nlhs > 3) plhs[3] = mask; else mxDestroyArray(mask);
    } else if (!strcmp(op, "wbbp")) {
        /* host computation: no device context is needed */
        if (nrhs != 8 || !mxIsChar(prhs[2]))
            mexErrMsgIdAndTxt("pkdecon:invalidArg", "wbbp takes psf, bp_type, alpha, beta, n, resFlag, iRes");
        static const char *types[] = {"traditional", "gaussian", "butterworth", "wiener", "wiener-butterworth"};
        char *ty = mxArrayToString(prhs[2]);
        int32_t t = -1;
        for (int i = 0; i < 5; ++i)
            if (!strcmp(ty, types[i])) t = i;
        mxFree(ty);
        if (t < 0) mexErrMsgIdAndTxt("pkdecon:invalidArg", "bp_type does not match any back-projector type");
        int64_t pd[3];
        dims3(prhs[1], pd);
        double ires[3] = {0, 0, 0};
        const size_t ni = mxGetNumberOfElements(prhs[7]);
        for (int i = 0; i < 3 && i < (int)ni; ++i) ires[i] = mxGetPr(prhs[7])[i];
        mwSize dm[3] = {(mwSize)pd[0], (mwSize)pd[1], (mwSize)pd[2]};
        plhs[0] = mxCreateNumericArray(3, dm, mxDOUBLE_CLASS, mxREAL);
        mxArray *otf = mxCreateNumericArray(3, dm, mxDOUBLE_CLASS, mxCOMPLEX);
        check(pk_wb_backprojector(NULL, single_arg(prhs[1], "psf"), pd, t, mxGetScalar(prhs[3]), mxGetScalar(prhs[4]),
                                  (int32_t)mxGetScalar(prhs[5]), (int32_t)mxGetScalar(prhs[6]), ires,
                                  (double *)mxGetData(plhs[0]), (double *)mxGetData(otf), NULL));
        if (nlhs > 1) plhs[1] = otf; else mxDestroyArray(otf);
    } else {
        mexErrMsgIdAndTxt("pkdecon:invalidArg", "unknown operation '%s'", op);
    }
    mxFree(op);
}
