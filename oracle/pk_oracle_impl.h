/*
 * pk_oracle_impl.h -- type-generic body of the oracle (TEST INFRASTRUCTURE).
 * Included twice by pk_oracle.c with REAL = float / double and SUF = f32 / f64.
 *
 * The FFT here is a plain mixed-radix Stockham transform over complex data,
 * the same full-spectrum arithmetic MATLAB's fftn/ifftn performs; it shares
 * no code with the HIP half-spectrum transforms it checks.
 */
#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, SUF)

typedef struct {
    int n, nf;
    int fac[64];
    REAL *twr, *twi; /* W_n^e = exp(-2 pi i e / n), e in [0, n) */
} F(plan1);

static F(plan1) *F(plan_create)(int n)
{
    F(plan1) *pl = (F(plan1) *)calloc(1, sizeof(*pl));
    if (!pl) return NULL;
    pl->n = n;
    int m = n;
    while (m % 4 == 0) { pl->fac[pl->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->fac[pl->nf++] = 2; m /= 2; }
    for (int p = 3; p * p <= m; p += 2)
        while (m % p == 0) { pl->fac[pl->nf++] = p; m /= p; }
    if (m > 1) pl->fac[pl->nf++] = m;
    pl->twr = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    pl->twi = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    for (int e = 0; e < n; e++) {
        /* computed in long double, rounded once to REAL */
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)e / (long double)n;
        pl->twr[e] = (REAL)cosl(a);
        pl->twi[e] = (REAL)(-sinl(a));
    }
    /* exact values at quarter turns */
    if (n % 4 == 0) { pl->twr[n / 4] = 0; pl->twi[n / 4] = -1; pl->twr[3 * n / 4] = 0; pl->twi[3 * n / 4] = 1; }
    if (n % 2 == 0) { pl->twr[n / 2] = -1; pl->twi[n / 2] = 0; }
    return pl;
}

static void F(plan_destroy)(F(plan1) *pl)
{
    if (!pl) return;
    free(pl->twr);
    free(pl->twi);
    free(pl);
}

/* One 1-D transform of n complex values (interleaved) in x, using scratch
 * (2n REAL).  Stockham autosort: stage with radix p after Ns = product of the
 * earlier radices reads x[j + q*n/p], twiddles by W_{Ns p}^{q k} (k = j mod Ns),
 * applies a p-point DFT and writes y[(j/Ns)*Ns*p + k + q*Ns]. */
static void F(fft1)(const F(plan1) *pl, REAL *x, REAL *scratch, int inverse)
{
    const int n = pl->n;
    if (n == 1) return;
    REAL *src = x, *dst = scratch;
    const REAL sg = inverse ? (REAL)1 : (REAL)-1; /* sign of the imaginary unit in W */
    int Ns = 1;
    for (int f = 0; f < pl->nf; f++) {
        const int p = pl->fac[f];
        const int m = n / p;
        const int tws = n / (Ns * p);
        REAL ar[p > 0 ? p : 1], ai[p > 0 ? p : 1];
        for (int j = 0; j < m; j++) {
            const int k = j % Ns;
            const int g = j / Ns;
            for (int q = 0; q < p; q++) {
                REAL xr = src[2 * (j + q * m)], xi = src[2 * (j + q * m) + 1];
                if (q && k) {
                    const int e = q * k * tws;
                    const REAL wr = pl->twr[e], wi = inverse ? -pl->twi[e] : pl->twi[e];
                    const REAL tr = xr * wr - xi * wi;
                    xi = xr * wi + xi * wr;
                    xr = tr;
                }
                ar[q] = xr;
                ai[q] = xi;
            }
            const int ob = g * Ns * p + k;
            if (p == 2) {
                dst[2 * ob] = ar[0] + ar[1];
                dst[2 * ob + 1] = ai[0] + ai[1];
                dst[2 * (ob + Ns)] = ar[0] - ar[1];
                dst[2 * (ob + Ns) + 1] = ai[0] - ai[1];
            } else if (p == 4) {
                const REAL s0r = ar[0] + ar[2], s0i = ai[0] + ai[2];
                const REAL d0r = ar[0] - ar[2], d0i = ai[0] - ai[2];
                const REAL s1r = ar[1] + ar[3], s1i = ai[1] + ai[3];
                const REAL d1r = ar[1] - ar[3], d1i = ai[1] - ai[3];
                /* sg*i*d1 = (-sg*d1i, sg*d1r) */
                dst[2 * ob] = s0r + s1r;
                dst[2 * ob + 1] = s0i + s1i;
                dst[2 * (ob + Ns)] = d0r - sg * d1i;
                dst[2 * (ob + Ns) + 1] = d0i + sg * d1r;
                dst[2 * (ob + 2 * Ns)] = s0r - s1r;
                dst[2 * (ob + 2 * Ns) + 1] = s0i - s1i;
                dst[2 * (ob + 3 * Ns)] = d0r + sg * d1i;
                dst[2 * (ob + 3 * Ns) + 1] = d0i - sg * d1r;
            } else {
                const int st = n / p;
                for (int q = 0; q < p; q++) {
                    REAL br = 0, bi = 0;
                    for (int mm = 0; mm < p; mm++) {
                        const int e = (int)(((long)mm * q) % p) * st;
                        const REAL wr = pl->twr[e], wi = inverse ? -pl->twi[e] : pl->twi[e];
                        br += ar[mm] * wr - ai[mm] * wi;
                        bi += ar[mm] * wi + ai[mm] * wr;
                    }
                    dst[2 * (ob + q * Ns)] = br;
                    dst[2 * (ob + q * Ns) + 1] = bi;
                }
            }
        }
        REAL *t = src;
        src = dst;
        dst = t;
        Ns *= p;
    }
    if (src != x) memcpy(x, src, sizeof(REAL) * 2 * (size_t)n);
}

#define PKO_BLK 8

int F(pko_fftn)(REAL *data, int64_t n1, int64_t n2, int64_t n3, int inverse)
{
    if (!data || n1 < 1 || n2 < 1 || n3 < 1) return PKO_ERR_ARG;
    F(plan1) *p1 = F(plan_create)((int)n1);
    F(plan1) *p2 = F(plan_create)((int)n2);
    F(plan1) *p3 = F(plan_create)((int)n3);
    if (!p1 || !p2 || !p3) { F(plan_destroy)(p1); F(plan_destroy)(p2); F(plan_destroy)(p3); return PKO_ERR_NOMEM; }
    const int64_t nmax = n1 > n2 ? (n1 > n3 ? n1 : n3) : (n2 > n3 ? n2 : n3);
    const int64_t plane = n1 * n2;
    int err = 0;
#pragma omp parallel
    {
        REAL *scr = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)nmax);
        REAL *col = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)nmax * PKO_BLK);
        if (!scr || !col) {
#pragma omp atomic write
            err = 1;
        }
        if (scr && col) {
            /* dim 1: contiguous rows */
            if (n1 > 1) {
#pragma omp for schedule(static)
                for (int64_t r = 0; r < n2 * n3; r++) F(fft1)(p1, data + 2 * r * n1, scr, inverse);
            }
            /* dim 2: blocks of PKO_BLK adjacent columns per z-slice */
            if (n2 > 1) {
                const int64_t nb1 = (n1 + PKO_BLK - 1) / PKO_BLK;
#pragma omp for schedule(static)
                for (int64_t t = 0; t < nb1 * n3; t++) {
                    const int64_t i3 = t / nb1, i1s = (t % nb1) * PKO_BLK;
                    const int bw = (int)((n1 - i1s) < PKO_BLK ? (n1 - i1s) : PKO_BLK);
                    REAL *base = data + 2 * (i1s + plane * i3);
                    for (int64_t i2 = 0; i2 < n2; i2++)
                        for (int b = 0; b < bw; b++) {
                            col[2 * (b * n2 + i2)] = base[2 * (b + n1 * i2)];
                            col[2 * (b * n2 + i2) + 1] = base[2 * (b + n1 * i2) + 1];
                        }
                    for (int b = 0; b < bw; b++) F(fft1)(p2, col + 2 * b * n2, scr, inverse);
                    for (int64_t i2 = 0; i2 < n2; i2++)
                        for (int b = 0; b < bw; b++) {
                            base[2 * (b + n1 * i2)] = col[2 * (b * n2 + i2)];
                            base[2 * (b + n1 * i2) + 1] = col[2 * (b * n2 + i2) + 1];
                        }
                }
            }
            /* dim 3: blocks of PKO_BLK adjacent (i1,i2) columns */
            if (n3 > 1) {
                const int64_t nb = (plane + PKO_BLK - 1) / PKO_BLK;
#pragma omp for schedule(static)
                for (int64_t t = 0; t < nb; t++) {
                    const int64_t cs = t * PKO_BLK;
                    const int bw = (int)((plane - cs) < PKO_BLK ? (plane - cs) : PKO_BLK);
                    REAL *base = data + 2 * cs;
                    for (int64_t i3 = 0; i3 < n3; i3++)
                        for (int b = 0; b < bw; b++) {
                            col[2 * (b * n3 + i3)] = base[2 * (b + plane * i3)];
                            col[2 * (b * n3 + i3) + 1] = base[2 * (b + plane * i3) + 1];
                        }
                    for (int b = 0; b < bw; b++) F(fft1)(p3, col + 2 * b * n3, scr, inverse);
                    for (int64_t i3 = 0; i3 < n3; i3++)
                        for (int b = 0; b < bw; b++) {
                            base[2 * (b + plane * i3)] = col[2 * (b * n3 + i3)];
                            base[2 * (b + plane * i3) + 1] = col[2 * (b * n3 + i3) + 1];
                        }
                }
            }
            if (inverse) {
                const REAL sc = (REAL)(1.0 / ((double)n1 * (double)n2 * (double)n3));
#pragma omp for schedule(static)
                for (int64_t i = 0; i < 2 * n1 * n2 * n3; i++) data[i] *= sc;
            }
        }
        free(scr);
        free(col);
    }
    F(plan_destroy)(p1);
    F(plan_destroy)(p2);
    F(plan_destroy)(p3);
    return err ? PKO_ERR_NOMEM : PKO_OK;
}

/* decon_psf2otf.m:1-46: pad post to outSize, circshift by -floor(psfSize/2),
 * fftn.  The imaginary part is kept (round-off discard is commented out in
 * the reference, :38-41); an all-zero PSF yields a zero OTF (:42-44). */
int F(pko_psf2otf)(const REAL *psf, const int64_t pd[3], const int64_t od[3], REAL *otf)
{
    if (!psf || !otf) return PKO_ERR_ARG;
    for (int d = 0; d < 3; d++)
        if (od[d] < pd[d]) return PKO_ERR_OUTSIZE;
    const int64_t N = od[0] * od[1] * od[2];
    memset(otf, 0, sizeof(REAL) * 2 * (size_t)N);
    int allzero = 1;
    const int64_t P = pd[0] * pd[1] * pd[2];
    for (int64_t i = 0; i < P; i++)
        if (psf[i] != 0) { allzero = 0; break; }
    if (allzero) return PKO_OK;
    const int64_t h0 = pd[0] / 2, h1 = pd[1] / 2, h2 = pd[2] / 2;
    for (int64_t k = 0; k < pd[2]; k++) {
        const int64_t z = ((k - h2) % od[2] + od[2]) % od[2];
        for (int64_t j = 0; j < pd[1]; j++) {
            const int64_t y = ((j - h1) % od[1] + od[1]) % od[1];
            for (int64_t i = 0; i < pd[0]; i++) {
                const int64_t x = ((i - h0) % od[0] + od[0]) % od[0];
                otf[2 * (x + od[0] * (y + od[1] * z))] = psf[i + pd[0] * (j + pd[1] * k)];
            }
        }
    }
    return F(pko_fftn)(otf, od[0], od[1], od[2], 0);
}

/* PSF ./ sum(PSF(:)) -- sum accumulated in double, rounded to REAL. */
static REAL *F(normalized_psf)(const REAL *psf, int64_t P)
{
    REAL *q = (REAL *)malloc(sizeof(REAL) * (size_t)P);
    if (!q) return NULL;
    double s = 0;
    for (int64_t i = 0; i < P; i++) s += (double)psf[i];
    const REAL sr = (REAL)s;
    for (int64_t i = 0; i < P; i++) q[i] = psf[i] / sr;
    return q;
}

static REAL *F(make_otf)(const REAL *psf, const int64_t pd[3], const int64_t od[3], int *err)
{
    const int64_t P = pd[0] * pd[1] * pd[2];
    const int64_t N = od[0] * od[1] * od[2];
    REAL *q = F(normalized_psf)(psf, P);
    REAL *otf = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)N);
    if (!q || !otf) { free(q); free(otf); *err = PKO_ERR_NOMEM; return NULL; }
    *err = F(pko_psf2otf)(q, pd, od, otf);
    free(q);
    if (*err) { free(otf); return NULL; }
    return otf;
}

/* out = max(real(ifftn(fftn(in) .* H)), lo) with H conjugated if conjH.
 * w is an N-complex work buffer. */
static int F(convolve)(const REAL *in, const REAL *H, int conjH, int64_t n1, int64_t n2, int64_t n3,
                       REAL *w, REAL *out_real)
{
    const int64_t N = n1 * n2 * n3;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) { w[2 * i] = in[i]; w[2 * i + 1] = 0; }
    int e = F(pko_fftn)(w, n1, n2, n3, 0);
    if (e) return e;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) {
        const REAL ar = w[2 * i], ai = w[2 * i + 1];
        const REAL hr = H[2 * i], hi = conjH ? -H[2 * i + 1] : H[2 * i + 1];
        w[2 * i] = ar * hr - ai * hi;
        w[2 * i + 1] = ar * hi + ai * hr;
    }
    e = F(pko_fftn)(w, n1, n2, n3, 1);
    if (e) return e;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) out_real[i] = w[2 * i];
    return PKO_OK;
}

/* Prep (decon_lucy_function.m:101-108, CPU branch): mask = I~=0 (only if
 * needed), I = max(I - Background, 0). */
static void F(prep)(const REAL *Iin, int64_t N, const pko_opts *o, REAL *I, uint8_t *mask)
{
    const REAL bg = (REAL)o->background;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) {
        if (mask) mask[i] = Iin[i] != 0;
        const REAL v = Iin[i] - bg;
        I[i] = v > 0 ? v : (REAL)0; /* max(x, 0); NaN -> 0 like MATLAB max */
    }
}

/* Epilogue (decon_lucy_function.m:177-197). */
static int F(epilogue)(REAL *J, const REAL *I, uint8_t *mask, const int64_t d[3], const pko_opts *o,
                       uint16_t *u16)
{
    const int64_t N = d[0] * d[1] * d[2];
    if (o->dampFactor > 1) {
        const REAL df = (REAL)o->dampFactor;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; i++) {
            const REAL dI = df * I[i];
            const REAL j = J[i];
            /* J.*(J<=dI) + min(J, dI.*(J>=dI)); equality gives 2J (reference quirk) */
            const REAL a = (j <= dI) ? j : (REAL)0;
            const REAL b2 = (j >= dI) ? dI : (REAL)0;
            const REAL b = j < b2 ? j : b2;
            J[i] = a + b;
        }
    }
    if (o->edgeErosion > 0 && mask) {
        int e = pko_mask_edge_erosion(mask, d, o->edgeErosion);
        if (e) return e;
    }
    if (o->scaleFactor != 1 || o->deconOffset != 0 || o->edgeErosion > 0) {
        const REAL sf = (REAL)o->scaleFactor, off = (REAL)o->deconOffset;
        const int use_mask = (o->deconOffset != 0 || o->edgeErosion > 0);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; i++) {
            if (use_mask) J[i] = (J[i] * sf + off) * (REAL)(mask ? mask[i] : 1);
            else J[i] = J[i] * sf;
        }
    }
    if (o->save16bit && u16) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; i++) u16[i] = pko_to_u16((double)J[i]);
    }
    return PKO_OK;
}

/* decon_lucy_function.m:57-213, CPU branch. */
int F(pko_decon_lucy)(const REAL *Iin, const int64_t d[3], const REAL *psf, const int64_t pd[3], int numit,
                      const pko_opts *o, REAL *J_out, uint16_t *u16)
{
    if (!Iin || !psf || !J_out || !o || numit < 0) return PKO_ERR_ARG;
    const int64_t N = d[0] * d[1] * d[2];
    int err = 0;
    REAL *H = F(make_otf)(psf, pd, d, &err);
    if (!H) return err;
    const int need_mask = (o->deconOffset != 0 || o->edgeErosion > 0);
    REAL *I = (REAL *)malloc(sizeof(REAL) * (size_t)N);
    REAL *J3 = (REAL *)calloc((size_t)N, sizeof(REAL));
    REAL *J4 = (REAL *)calloc((size_t)N, sizeof(REAL));
    REAL *Y = (REAL *)calloc((size_t)N, sizeof(REAL));
    REAL *R = (REAL *)malloc(sizeof(REAL) * (size_t)N);
    REAL *w = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)N);
    uint8_t *mask = need_mask ? (uint8_t *)malloc((size_t)N) : NULL;
    REAL *J2 = J_out;
    if (!I || !J3 || !J4 || !Y || !R || !w || (need_mask && !mask)) { err = PKO_ERR_NOMEM; goto done; }
    F(prep)(Iin, N, o, I, mask);
    memcpy(J2, I, sizeof(REAL) * (size_t)N);
    {
        const REAL eps = (REAL)PKO_EPS;
        REAL lambda = 0; /* 2*any(J_4(:)~=0) with J_4 = 0 (:120) */
        for (int k = 1; k <= numit; k++) {
            if (k > 2) {
                double num = 0, den = 0;
#pragma omp parallel for schedule(static) reduction(+ : num, den)
                for (int64_t i = 0; i < N; i++) {
                    num += (double)((J2[i] - Y[i]) * J4[i]);
                    den += (double)(J4[i] * J4[i]);
                }
                lambda = (REAL)num / ((REAL)den + eps);
                lambda = lambda > 1 ? (REAL)1 : lambda; /* max(min(lambda,1),0) */
                lambda = lambda > 0 ? lambda : (REAL)0;
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < N; i++) J4[i] = J2[i] - Y[i];
            } else if (k == 2) {
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < N; i++) J4[i] = J2[i] - Y[i];
            }
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; i++) {
                const REAL y = J2[i] + lambda * (J2[i] - J3[i]);
                Y[i] = y > 0 ? y : (REAL)0;
            }
            /* ReBlurred = max(real(ifftn(H.*fftn(Y))), eps); ReBlurred = I./ReBlurred (:136-138) */
            if ((err = F(convolve)(Y, H, 0, d[0], d[1], d[2], w, R))) goto done;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; i++) {
                const REAL rb = R[i] > eps ? R[i] : eps;
                R[i] = I[i] / rb;
            }
            /* J_3 = J_2; J_2 = max(Y.*real(ifftn(conj(H).*fftn(ReBlurred))), 0) (:141-142) */
            if ((err = F(convolve)(R, H, 1, d[0], d[1], d[2], w, R))) goto done;
            memcpy(J3, J2, sizeof(REAL) * (size_t)N);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; i++) {
                const REAL v = Y[i] * R[i];
                J2[i] = v > 0 ? v : (REAL)0;
            }
        }
    }
    err = F(epilogue)(J2, I, mask, d, o, u16);
done:
    free(H); free(I); free(J3); free(J4); free(Y); free(R); free(w); free(mask);
    return err;
}

/* decon_lucy_omw_function.m:79-178 (prep :90-98, loop :105-112, epilogue :145-176), CPU branch. */
int F(pko_decon_omw)(const REAL *Iin, const int64_t d[3], const REAL *psf_f, const int64_t pfd[3],
                     const REAL *psf_b, const int64_t pbd[3], int numit, const pko_opts *o, REAL *J_out,
                     uint16_t *u16)
{
    if (!Iin || !psf_f || !psf_b || !J_out || !o || numit < 0) return PKO_ERR_ARG;
    const int64_t N = d[0] * d[1] * d[2];
    int err = 0;
    REAL *Hf = F(make_otf)(psf_f, pfd, d, &err);
    if (!Hf) return err;
    REAL *Hb = F(make_otf)(psf_b, pbd, d, &err);
    if (!Hb) { free(Hf); return err; }
    const int need_mask = (o->deconOffset != 0 || o->edgeErosion > 0);
    REAL *I = (REAL *)malloc(sizeof(REAL) * (size_t)N);
    REAL *C = (REAL *)malloc(sizeof(REAL) * (size_t)N);
    REAL *w = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)N);
    uint8_t *mask = need_mask ? (uint8_t *)malloc((size_t)N) : NULL;
    REAL *J = J_out;
    if (!I || !C || !w || (need_mask && !mask)) { err = PKO_ERR_NOMEM; goto done; }
    F(prep)(Iin, N, o, I, mask);
    memcpy(J, I, sizeof(REAL) * (size_t)N);
    {
        const REAL eps = (REAL)PKO_EPS;
        for (int k = 1; k <= numit; k++) {
            /* CX = max(real(ifftn(fftn(J).*OTF_f)), eps) (:323-325) */
            if ((err = F(convolve)(J, Hf, 0, d[0], d[1], d[2], w, C))) goto done;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; i++) {
                const REAL c = C[i] > eps ? C[i] : eps;
                C[i] = I[i] / c;
            }
            /* J = max(real(ifftn(fftn(I./CX).*OTF_b)).*J, 0) (:326-328) */
            if ((err = F(convolve)(C, Hb, 0, d[0], d[1], d[2], w, C))) goto done;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; i++) {
                const REAL v = C[i] * J[i];
                J[i] = v > 0 ? v : (REAL)0;
            }
        }
    }
    err = F(epilogue)(J, I, mask, d, o, u16);
done:
    free(Hf); free(Hb); free(I); free(C); free(w); free(mask);
    return err;
}

#undef F
#undef CAT
#undef CAT2
