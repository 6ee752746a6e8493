/* Test harness: a tiny in-memory implementation of the mx API declared in
 * mex.h, plus a driver that calls mexFunction and reports errors. */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls;
    mwSize nd, dims[3];
    void *data;
    char *str;
    const char *fname[8];
    mxArray *fval[8];
    int nf;
};

static jmp_buf g_err;
static char g_errid[128];

/* MATLAB's memory manager frees the mxArrays and mxMalloc blocks a MEX call
 * leaves behind; this registry frees them at exit, so a leak checker run on
 * the harness reports only memory the gateway or the library itself lose. */
static void **g_blocks;
static size_t g_nblocks, g_capblocks;
static void *mx_alloc(size_t n) {
    void *p = calloc(1, n ? n : 1);
    if (g_nblocks == g_capblocks) {
        g_capblocks = g_capblocks ? 2 * g_capblocks : 256;
        g_blocks = realloc(g_blocks, g_capblocks * sizeof(void *));
    }
    g_blocks[g_nblocks++] = p;
    return p;
}
static void mx_release(void *p) {
    if (!p) return;
    for (size_t i = 0; i < g_nblocks; ++i)
        if (g_blocks[i] == p) { g_blocks[i] = g_blocks[--g_nblocks]; free(p); return; }
    free(p);
}
static void mx_release_all(void) {
    for (size_t i = 0; i < g_nblocks; ++i) free(g_blocks[i]);
    free(g_blocks);
    g_blocks = NULL;
    g_nblocks = g_capblocks = 0;
}
static char *mx_strdup(const char *s) { char *d = mx_alloc(strlen(s) + 1); strcpy(d, s); return d; }

static size_t esz(mxClassID c) { return c == mxDOUBLE_CLASS ? 8 : c == mxSINGLE_CLASS ? 4 : c == mxUINT16_CLASS ? 2 : 1; }
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->nd; }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims; }
size_t mxGetNumberOfElements(const mxArray *a) { size_t n = 1; for (mwSize i = 0; i < a->nd; ++i) n *= a->dims[i]; return n; }
mxClassID mxGetClassID(const mxArray *a) { return a->cls; }
int mxIsStruct(const mxArray *a) { return a && a->cls == mxSTRUCT_CLASS; }
int mxIsChar(const mxArray *a) { return a->cls == mxCHAR_CLASS; }
int mxIsSingle(const mxArray *a) { return a->cls == mxSINGLE_CLASS; }
int mxIsEmpty(const mxArray *a) { return mxGetNumberOfElements(a) == 0; }
mxArray *mxGetField(const mxArray *a, size_t i, const char *name) {
    (void)i;
    for (int k = 0; k < a->nf; ++k) if (!strcmp(a->fname[k], name)) return a->fval[k];
    return NULL;
}
double mxGetScalar(const mxArray *a) { return a->cls == mxDOUBLE_CLASS ? ((double *)a->data)[0] : 0.0; }
double *mxGetPr(const mxArray *a) { return (double *)a->data; }
void *mxGetData(const mxArray *a) { return a->data; }
char *mxArrayToString(const mxArray *a) { return mx_strdup(a->str); }
void *mxMalloc(size_t n) { return mx_alloc(n); }
void mxFree(void *p) { mx_release(p); }
mxArray *mxCreateNumericArray(mwSize nd, const mwSize *dims, mxClassID c, mxComplexity cx) {
    mxArray *a = mx_alloc(sizeof(*a));
    a->cls = c; a->nd = nd;
    for (mwSize i = 0; i < nd; ++i) a->dims[i] = dims[i];
    a->data = mx_alloc((mxGetNumberOfElements(a) ? mxGetNumberOfElements(a) : 1) * esz(c) * (cx == mxCOMPLEX ? 2 : 1));
    return a;
}
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity cx) { mwSize d[2] = {m, n}; return mxCreateNumericArray(2, d, mxDOUBLE_CLASS, cx); }
mxArray *mxCreateDoubleScalar(double v) { mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL); ((double *)a->data)[0] = v; return a; }
void mxDestroyArray(mxArray *a) { if (a) { mx_release(a->data); mx_release(a->str); mx_release(a); } }
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
    (void)fmt;
    snprintf(g_errid, sizeof(g_errid), "%s", id);
    longjmp(g_err, 1);
}
int mexAtExit(void (*fn)(void)) { (void)fn; return 0; }

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

static mxArray *str(const char *s) { mxArray *a = mx_alloc(sizeof(*a)); a->cls = mxCHAR_CLASS; a->nd = 2; a->dims[0] = 1; a->dims[1] = strlen(s); a->str = mx_strdup(s); return a; }
static mxArray *dbl(double v) { return mxCreateDoubleScalar(v); }

static int call(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs) {
    if (setjmp(g_err)) { printf("error %s\n", g_errid); return 1; }
    mexFunction(nlhs, plhs, nrhs, prhs);
    return 0;
}

static void *slurp(const char *dir, const char *name, size_t bytes) {
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) { printf("missing %s\n", path); exit(2); }
    void *p = malloc(bytes);
    if (fread(p, 1, bytes, f) != bytes) { printf("short %s\n", path); exit(2); }
    fclose(f);
    return p;
}

static void spill(const char *dir, const char *name, const void *p, size_t bytes) {
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(p, 1, bytes, f) != bytes) { printf("cannot write %s\n", path); exit(2); }
    fclose(f);
}

static mxArray *opts_struct(double background, double save16bit) {
    mxArray *s = mx_alloc(sizeof(*s));
    s->cls = mxSTRUCT_CLASS; s->nd = 2; s->dims[0] = s->dims[1] = 1;
    s->fname[0] = "Background"; s->fval[0] = dbl(background);
    s->fname[1] = "save16bit"; s->fval[1] = dbl(save16bit);
    s->nf = 2;
    return s;
}

/* files <dir> n1 n2 n3 p1 p2 p3 numit_rl numit_omw background: reads I.u16,
 * psf.f32 and psfb.f32 from dir, runs 'rl' and 'omw' through mexFunction with
 * single outputs, writes J_rl.f32 and J_omw.f32 (column-major) and
 * err_rl.f64 (the err_mat output). */
static int files_mode(int argc, char **argv) {
    if (argc != 12) { printf("usage: files dir n1 n2 n3 p1 p2 p3 numit_rl numit_omw background\n"); return 2; }
    const char *dir = argv[2];
    mwSize d[3], pd[3];
    for (int i = 0; i < 3; ++i) { d[i] = (mwSize)atol(argv[3 + i]); pd[i] = (mwSize)atol(argv[6 + i]); }
    const double it_rl = atof(argv[9]), it_omw = atof(argv[10]), bg = atof(argv[11]);
    const size_t n = (size_t)d[0] * d[1] * d[2], np = (size_t)pd[0] * pd[1] * pd[2];
    mxArray *I = mxCreateNumericArray(3, d, mxUINT16_CLASS, mxREAL);
    mxArray *P = mxCreateNumericArray(3, pd, mxSINGLE_CLASS, mxREAL);
    mxArray *PB = mxCreateNumericArray(3, pd, mxSINGLE_CLASS, mxREAL);
    void *t = slurp(dir, "I.u16", n * 2); memcpy(I->data, t, n * 2); free(t);
    t = slurp(dir, "psf.f32", np * 4); memcpy(P->data, t, np * 4); free(t);
    t = slurp(dir, "psfb.f32", np * 4); memcpy(PB->data, t, np * 4); free(t);
    mxArray *s = opts_struct(bg, 0);
    mxArray *out[4] = {0};
    const mxArray *a1[] = {str("rl"), I, P, dbl(it_rl), s};
    if (call(2, out, 5, a1)) return 1;
    spill(dir, "J_rl.f32", out[0]->data, n * 4);
    spill(dir, "err_rl.f64", out[1]->data, (size_t)it_rl * 4 * 8);
    printf("rl ok %zu %zu %zu\n", out[0]->dims[0], out[0]->dims[1], out[0]->dims[2]);
    const mxArray *a2[] = {str("omw"), I, P, PB, dbl(it_omw), s};
    if (call(1, out, 6, a2)) return 1;
    spill(dir, "J_omw.f32", out[0]->data, n * 4);
    printf("omw ok %zu %zu %zu\n", out[0]->dims[0], out[0]->dims[1], out[0]->dims[2]);
    const mxArray *a3[] = {str("reset")};
    if (call(0, out, 1, a3)) return 1;
    printf("reset ok\n");
    return 0;
}

int main(int argc, char **argv) {
    atexit(mx_release_all);
    if (argc > 1 && !strcmp(argv[1], "files")) return files_mode(argc, argv);
    mxArray *out[4] = {0};
    const mxArray *a1[] = {str("fgfn"), dbl(101), dbl(1), dbl(0), dbl(7)};
    if (!call(1, out, 5, a1)) printf("fgfn %g\n", ((double *)out[0]->data)[0]);
    const mxArray *a2[] = {str("fgfn"), dbl(257), dbl(1), dbl(0), dbl(5)};
    if (!call(1, out, 5, a2)) printf("fgfn %g\n", ((double *)out[0]->data)[0]);
    const mxArray *a3[] = {str("nosuchop")};
    call(0, out, 1, a3);
    mwSize d[3] = {8, 8, 4};
    mxArray *I = mxCreateNumericArray(3, d, mxDOUBLE_CLASS, mxREAL);
    mwSize pdm[3] = {3, 3, 3};
    mxArray *P = mxCreateNumericArray(3, pdm, mxSINGLE_CLASS, mxREAL);
    mxArray *s = mx_alloc(sizeof(*s)); s->cls = mxSTRUCT_CLASS; s->nd = 2; s->dims[0] = s->dims[1] = 1;
    const mxArray *a4[] = {str("rl"), I, P, dbl(2), s};
    call(1, out, 5, a4);  /* double input -> unsupportedType */
    mxArray *I16 = mxCreateNumericArray(3, d, mxUINT16_CLASS, mxREAL);
    const mxArray *a5[] = {str("rl"), I16, P, dbl(2), s};
    if (!call(2, out, 5, a5)) printf("rl ok %zu %zu %zu\n", out[0]->dims[0], out[0]->dims[1], out[0]->dims[2]);
    /* BackProjector (host): delta PSF, wiener-butterworth; then an unknown type */
    mwSize wd[3] = {9, 9, 9};
    mxArray *D = mxCreateNumericArray(3, wd, mxSINGLE_CLASS, mxREAL);
    ((float *)D->data)[4 + 9 * (4 + 9 * 4)] = 1.f;
    mwSize rd[2] = {1, 3};
    mxArray *ir = mxCreateNumericArray(2, rd, mxDOUBLE_CLASS, mxREAL);
    const mxArray *a6[] = {str("wbbp"), D, str("wiener-butterworth"), dbl(0.005), dbl(1), dbl(10), dbl(1), ir};
    if (!call(2, out, 8, a6))
        printf("wbbp ok %zu %zu %zu %d\n", out[0]->dims[0], out[0]->dims[1], out[0]->dims[2],
               ((double *)out[0]->data)[4 + 9 * (4 + 9 * 4)] > 0);
    const mxArray *a7[] = {str("wbbp"), D, str("nosuchtype"), dbl(0.005), dbl(1), dbl(10), dbl(1), ir};
    call(2, out, 8, a7);
    return 0;
}
