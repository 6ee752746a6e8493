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

/* Binary erosion of every z-slice by a disk of radius r (offsets with
 * dy^2+dx^2 <= r^2), out-of-image neighbours counted as true (imerode's
 * padding for binary erosion). */
static void erode_disk2d(uint8_t *m, const int64_t d[3], int r)
{
    if (r <= 0) return;
    const int64_t n1 = d[0], n2 = d[1], n3 = d[2];
    uint8_t *t = (uint8_t *)malloc((size_t)(n1 * n2));
    for (int64_t z = 0; z < n3; z++) {
        uint8_t *s = m + z * n1 * n2;
        memcpy(t, s, (size_t)(n1 * n2));
        for (int64_t x = 0; x < n2; x++)
            for (int64_t y = 0; y < n1; y++) {
                uint8_t v = 1;
                for (int dx = -r; dx <= r && v; dx++)
                    for (int dy = -r; dy <= r && v; dy++) {
                        if (dx * dx + dy * dy > r * r) continue;
                        const int64_t yy = y + dy, xx = x + dx;
                        if (yy < 0 || yy >= n1 || xx < 0 || xx >= n2) continue;
                        v = t[yy + n1 * xx];
                    }
                s[y + n1 * x] = v;
            }
    }
    free(t);
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

/* decon_mask_edge_erosion.m:1-30.  The irregular-mask branch uses an exact
 * disk for strel('disk', E-1) (MATLAB uses the exact disk for radius < 3 and
 * a periodic-line approximation above). */
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
