/*
 * pk_oracle.c -- CPU restatement of PetaKit5D's RL / OMW deconvolution path.
 * TEST INFRASTRUCTURE ONLY (see pk_oracle.h): the checker, never the product.
 */
#include "pk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* MATLAB uint16(x): round to nearest, ties away from zero; saturate; NaN -> 0. */
uint16_t pko_to_u16(double x)
{
    if (!(x > 0)) return 0; /* also NaN */
    if (x >= 65535.0) return 65535;
    return (uint16_t)round(x);
}

static const double PKO_EPS = 2.220446049250313080847e-16; /* MATLAB eps = 2^-52 */

#define REAL float
#define SUF f32
#include "pk_oracle_impl.h"
#undef REAL
#undef SUF

#define REAL double
#define SUF f64
#include "pk_oracle_impl.h"
#undef REAL
#undef SUF

/* utils/findGoodFactorNumber.m:9-24: step by dir until the largest prime
 * factor is <= max_prime; with max_prime=7, odd numbers whose largest factor
 * is 7 are rejected unless allow_odd_seven (:21). */
int64_t pko_find_good_factor_number(int64_t n, int dir, int allow_odd_seven, int max_prime)
{
    if (dir == 0) dir = 1;
    for (;;) {
        if (n <= 1) return n; /* factor(1) = 1, factor(0) = 0: loop exits */
        int64_t m = n, pmin = 0, pmax = 1;
        for (int64_t p = 2; p * p <= m; p++)
            while (m % p == 0) {
                if (!pmin) pmin = p;
                pmax = p;
                m /= p;
            }
        if (m > 1) {
            if (!pmin) pmin = m;
            if (m > pmax) pmax = m;
        }
        int bad = pmax > max_prime;
        if (max_prime >= 7 && !allow_odd_seven && pmin > 2 && pmax == 7) bad = 1;
        if (!bad) return n;
        n += dir;
    }
}

/* Neighbourhood of MATLAB's strel('disk', r) with the default N = 4
 * (decon_mask_edge_erosion.m:22-27 calls strel('disk', EdgeErosion - 1)).
 * Image Processing Toolbox, as documented for strel: for r < 3 the disk is
 * exact (dx^2 + dy^2 <= r^2); for r >= 3 it is the radial decomposition into
 * periodic lines of Adams (CVGIP: Graphical Models and Image Processing 55(5),
 * 1993) with the periodic lines of Jones & Soille (Pattern Recognition Letters
 * 17, 1996): offsets v = (1,0), (1,1), (0,1), (-1,1) (row, column), each
 * repeated rp = floor(k / |v|) times either side with k = 2r / (cot(t) +
 * 1/sin(t)), t = pi / (2N); the Minkowski sum of those lines falls short of
 * radius r, so a horizontal and a vertical line of length 2(r - e - 1) + 1 are
 * added when that is >= 3, e being the sum's extent.  The result is convex and
 * point-symmetric: every row dy in [-R, R] is the run |dx| <= w[dy + R].
 * Returns R (w needs 2r + 1 entries), or -1 for r < 0. */
int pko_strel_disk(int r, int *w)
{
    if (r < 0) return -1;
    if (r < 3) {
        for (int dy = -r; dy <= r; dy++) {
            int h = 0;
            while ((h + 1) * (h + 1) + dy * dy <= r * r) h++;
            w[dy + r] = h;
        }
        return r;
    }
    const int v[4][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}};
    const double t = 3.14159265358979323846 / 8.0;
    const double k = 2.0 * r / (1.0 / tan(t) + 1.0 / sin(t));
    int rp[4];
    for (int q = 0; q < 4; q++) rp[q] = (int)floor(k / sqrt((double)(v[q][0] * v[q][0] + v[q][1] * v[q][1])));
    /* Minkowski sum on a (2r+1)^2 grid (the sum never reaches past r) */
    const int G = 2 * r + 1;
    uint8_t *a = (uint8_t *)calloc((size_t)G * G, 1), *b = (uint8_t *)malloc((size_t)G * G);
    a[r * G + r] = 1;
    for (int q = 0; q < 4; q++) {
        memset(b, 0, (size_t)G * G);
        for (int y = 0; y < G; y++)
            for (int x = 0; x < G; x++) {
                if (!a[y * G + x]) continue;
                for (int i = -rp[q]; i <= rp[q]; i++) {
                    const int yy = y + i * v[q][0], xx = x + i * v[q][1];
                    if (yy >= 0 && yy < G && xx >= 0 && xx < G) b[yy * G + xx] = 1;
                }
            }
        memcpy(a, b, (size_t)G * G);
    }
    int e = 0; /* row extent of the sum */
    for (int y = 0; y < G; y++)
        for (int x = 0; x < G; x++)
            if (a[y * G + x] && abs(y - r) > e) e = abs(y - r);
    const int len = 2 * (r - e - 1) + 1;
    const int h = len >= 3 ? (len - 1) / 2 : 0; /* half-length of the added lines */
    const int R = e + h;
    for (int dy = -R; dy <= R; dy++) {
        int best = -1; /* widest run of the sum within h rows of dy, widened by h */
        for (int sy = dy - h; sy <= dy + h; sy++) {
            if (sy < -e || sy > e) continue;
            for (int x = 0; x < G; x++)
                if (a[(sy + r) * G + x] && abs(x - r) > best) best = abs(x - r);
        }
        w[dy + R] = best + h;
    }
    free(a);
    free(b);
    return R;
}

/* Binary erosion of every z-slice by strel('disk', r) (pko_strel_disk),
 * out-of-image neighbours counted as true (imerode's padding for binary
 * erosion, applied once around the whole decomposed element). */
static void erode_disk2d(uint8_t *m, const int64_t d[3], int r)
{
    if (r <= 0) return;
    int *w = (int *)malloc(sizeof(int) * (size_t)(2 * r + 1));
    const int R = pko_strel_disk(r, w);
    const int64_t n1 = d[0], n2 = d[1], n3 = d[2];
    uint8_t *t = (uint8_t *)malloc((size_t)(n1 * n2));
    for (int64_t z = 0; z < n3; z++) {
        uint8_t *s = m + z * n1 * n2;
        memcpy(t, s, (size_t)(n1 * n2));
        for (int64_t x = 0; x < n2; x++)
            for (int64_t y = 0; y < n1; y++) {
                uint8_t v = 1;
                for (int dy = -R; dy <= R && v; dy++)
                    for (int dx = -w[dy + R]; dx <= w[dy + R] && v; dx++) {
                        const int64_t yy = y + dy, xx = x + dx;
                        if (yy < 0 || yy >= n1 || xx < 0 || xx >= n2) continue;
                        v = t[yy + n1 * xx];
                    }
                s[y + n1 * x] = v;
            }
    }
    free(t);
    free(w);
}

static void erode_line_z(uint8_t *m, const int64_t d[3], int half)
{
    if (half <= 0) return;
    const int64_t plane = d[0] * d[1], n3 = d[2];
    uint8_t *t = (uint8_t *)malloc((size_t)(plane * n3));
    memcpy(t, m, (size_t)(plane * n3));
    for (int64_t z = 0; z < n3; z++)
        for (int64_t i = 0; i < plane; i++) {
            uint8_t v = 1;
            for (int dz = -half; dz <= half && v; dz++) {
                const int64_t zz = z + dz;
                if (zz < 0 || zz >= n3) continue;
                v = t[i + plane * zz];
            }
            m[i + plane * z] = v;
        }
    free(t);
}

/* decon_mask_edge_erosion.m:1-30.  The irregular-mask branch erodes each
 * slice by strel('disk', E-1) (pko_strel_disk) and along dim 3 by a
 * (2E-1)-long line. */
int pko_mask_edge_erosion(uint8_t *mask, const int64_t d[3], int E)
{
    if (!mask || E < 0) return PKO_ERR_ARG;
    const int64_t n1 = d[0], n2 = d[1], n3 = d[2], N = n1 * n2 * n3;
    const int is3d = n3 > 1;
    int64_t cnt = 0;
    for (int64_t i = 0; i < N; i++) {
        mask[i] = mask[i] > 0;
        cnt += mask[i];
    }
    if (cnt == N) {
        for (int64_t z = 0; z < n3; z++)
            for (int64_t x = 0; x < n2; x++)
                for (int64_t y = 0; y < n1; y++) {
                    int clr = (y < E || y >= n1 - E || x < E || x >= n2 - E);
                    if (is3d && (z < E || z >= n3 - E)) clr = 1;
                    if (clr) mask[y + n1 * (x + n2 * z)] = 0;
                }
        return PKO_OK;
    }
    for (int64_t z = 0; z < n3; z++)
        for (int64_t x = 0; x < n2; x++)
            for (int64_t y = 0; y < n1; y++) {
                int clr = (y == 0 || y == n1 - 1 || x == 0 || x == n2 - 1);
                if (is3d && (z == 0 || z == n3 - 1)) clr = 1;
                if (clr) mask[y + n1 * (x + n2 * z)] = 0;
            }
    erode_disk2d(mask, d, E - 1);
    if (is3d) erode_line_z(mask, d, E - 1);
    return PKO_OK;
}
