// pk_fft_dev.hpp -- device side of the gfx950 deconvolution engine.
//
// Volumes are MATLAB column-major: dim 1 (n1) contiguous.  Real volumes hold
// N = n1*n2*n3 floats; spectra hold the half spectrum along dim 1,
// h1 = floor(n1/2)+1 complex values per row, S = h1*n2*n3 float2.
//
// One 1-D transform of length L is executed by P threads that each hold E
// values in registers (L = P*E).  Thread p always owns the "home" positions
// p + e*P (e in [0,E)).  A Stockham stage of radix R with Ns = product of the
// earlier radices reads positions j + m*L/R (j = p + i*P, i < E/R), which are
// exactly home positions, and writes (j/Ns)*Ns*R + (j mod Ns) + q*Ns; those
// go through LDS and are re-read at home positions by the next stage.  The
// last stage's outputs are home positions again, so the first stage of a
// transform and the last stage need no LDS round trip, and a forward
// transform, a pointwise product and an inverse transform chain in registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pk_dft.hpp"

namespace pk {

constexpr int kMaxStages = 10;
constexpr int kMaxThreads = 256;

struct FftPlan {
    int L;                      // transform length (complex points)
    int P;                      // threads per transform
    int E;                      // values per thread, L = P * E
    int nst;                    // number of Stockham stages
    int R[kMaxStages];          // radix per stage
    int Ns[kMaxStages];         // product of the radices of earlier stages
    uint32_t mag[kMaxStages];   // floor(2^32 / Ns) + 1: j / Ns = umulhi(j, mag)
    int tws[kMaxStages];        // twiddle table stride L / (Ns * R)
    const float2* tw;           // W_L^e = exp(-2 pi i e / L), e in [0, L)
};

// ------------------------------------------------------------------ one stage
template <int E, int R, bool INV>
__device__ __forceinline__ void stage_r(float2 (&v)[E], const FftPlan& pl, int s, int p, float2* lds, int lbase,
                                        int lstride, bool last) {
    constexpr int Q = E / R;  // butterflies per thread
    const int P = pl.P;
    const int Ns = pl.Ns[s];
    const int tws = pl.tws[s];
    const uint32_t mag = pl.mag[s];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int j = p + i * P;
        int g, k;
        if (s == 0) {
            g = j;
            k = 0;
        } else {
            g = (int)__umulhi((uint32_t)j, mag);
            k = j - g * Ns;
        }
        float2 u[R];
        static_for<0, R>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            u[m] = v[i + m * Q];
        });
        if (k != 0) {
            static_for<1, R>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                const float2 w = pl.tw[m * k * tws];
                u[m] = INV ? cmulc(u[m], w) : cmul(u[m], w);
            });
        }
        dft<R, INV>(u);
        if (last) {
            static_for<0, R>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                v[i + q * Q] = u[q];
            });
        } else {
            const int ob = g * Ns * R + k;
            static_for<0, R>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                lds[lbase + (ob + q * Ns) * lstride] = u[q];
            });
        }
    }
}

#define PK_RADIX_CASE(RR)                                                    \
    case RR:                                                                 \
        if constexpr (E % RR == 0) stage_r<E, RR, INV>(v, pl, s, p, lds, lbase, lstride, last); \
        break;

// Full transform of the E home values of every thread (all threads of the
// block must call it: it contains barriers).
template <int E, bool INV>
__device__ __forceinline__ void fft_run(float2 (&v)[E], const FftPlan& pl, int p, float2* lds, int lbase,
                                        int lstride) {
    for (int s = 0; s < pl.nst; ++s) {
        const bool last = (s == pl.nst - 1);
        switch (pl.R[s]) {
            PK_RADIX_CASE(2)
            PK_RADIX_CASE(3)
            PK_RADIX_CASE(4)
            PK_RADIX_CASE(5)
            PK_RADIX_CASE(6)
            PK_RADIX_CASE(8)
            PK_RADIX_CASE(9)
            PK_RADIX_CASE(10)
            PK_RADIX_CASE(12)
            PK_RADIX_CASE(15)
            PK_RADIX_CASE(16)
            PK_RADIX_CASE(18)
            PK_RADIX_CASE(20)
            PK_RADIX_CASE(24)
            PK_RADIX_CASE(25)
            PK_RADIX_CASE(27)
            PK_RADIX_CASE(30)
            PK_RADIX_CASE(32)
            default:
                break;
        }
        if (!last) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[lbase + (p + e * pl.P) * lstride];
            __syncthreads();
        }
    }
}
#undef PK_RADIX_CASE

// ------------------------------------------------------------------ x-axis (rows)
enum XMode : int {
    XM_R2C = 0,          // real src -> spectrum
    XM_C2R = 1,          // spectrum -> real dst (unnormalised)
    XM_PREP_U16 = 2,     // raw uint16: I = sat(raw - bg16); I -> dst_I, J; spectrum of I
    XM_PREP_F32 = 3,     // raw float: I = max(raw - bg, 0); I -> dst_I, J; spectrum of I
    XM_RATIO = 4,        // spectrum -> c; r = I / max(c*scale, eps); spectrum of r
    XM_OMW_UPDATE = 5,   // spectrum -> t; J = max(t*scale * J, 0); [spectrum of J]
    XM_BIGGS_Y = 6,      // Y = max(J2 + lambda (J2 - J3), 0) -> Y; spectrum of Y
    XM_BIGGS_UPDATE = 7, // spectrum -> t; Jn = max(Y * t*scale, 0) -> Jn; J4 = Jn - Y; dot partials
};

struct XArgs {
    int n1;            // real row length
    int h1;            // n1/2 + 1
    int packed;        // n1 even: transform length n1/2 over (x[2m], x[2m+1])
    int RT;            // rows per block
    int lpad;          // LDS row pitch (float2)
    int mode;
    int r2c_out;       // modes 5/7: also produce the spectrum of the new estimate
    int acc;           // mode 7: accumulate the two dot products
    int64_t nrows;     // n2 * n3
    float scale;       // 1/N folded into the inverse transform
    float bg;          // background (float) / bg16 for uint16
    float eps;         // MATLAB eps as single
    float2* spec;      // half spectrum (in and/or out)
    const float2* rtw; // W_{n1}^k, k in [0, n1/2]  (packed post-processing)
    const float* src;  // XM_R2C input
    float* dst;        // XM_C2R output
    const uint16_t* raw16;
    const float* rawf;
    float* I;          // background-subtracted input
    float* J;          // estimate (OMW J, Biggs J2 for mode 6)
    const float* J3;   // Biggs x_{k-2}
    float* Y;          // Biggs prediction
    float* Jn;         // Biggs new estimate
    float* J4;         // Biggs x_{k-1} - y_{k-1}
    const float* lambda;  // device scalar
    double* partials;     // 2 doubles per block
};

__device__ __forceinline__ float fmax0(float a) { return a > 0.f ? a : 0.f; }  // max(a,0), NaN -> 0

// Elementwise step in real space on one value x at real index gi.
// Returns the value that feeds the forward transform (if any).
__device__ __forceinline__ float x_elem(const XArgs& a, int64_t gi, float x, double& s_num, double& s_den) {
    switch (a.mode) {
        case XM_R2C:
            return a.src[gi];
        case XM_C2R:
            a.dst[gi] = x;
            return 0.f;
        case XM_PREP_U16: {
            const float r = a.raw16[gi];
            const float v = fmax0(r - a.bg);  // uint16 saturating subtract of an integer
            a.I[gi] = v;
            a.J[gi] = v;
            return v;
        }
        case XM_PREP_F32: {
            const float v = fmax0(a.rawf[gi] - a.bg);
            a.I[gi] = v;
            a.J[gi] = v;
            return v;
        }
        case XM_RATIO: {
            const float c = x * a.scale;
            const float cm = c > a.eps ? c : a.eps;  // max(CX, eps)
            return a.I[gi] / cm;
        }
        case XM_OMW_UPDATE: {
            const float v = fmax0((x * a.scale) * a.J[gi]);
            a.J[gi] = v;
            return v;
        }
        case XM_BIGGS_Y: {
            const float lam = *a.lambda;
            const float j2 = a.J[gi];
            const float y = fmax0(j2 + lam * (j2 - a.J3[gi]));
            a.Y[gi] = y;
            return y;
        }
        case XM_BIGGS_UPDATE: {
            const float y = a.Y[gi];
            const float v = fmax0(y * (x * a.scale));
            a.Jn[gi] = v;
            const float d = v - y;
            if (a.acc) {
                const float j4 = a.J4[gi];
                s_num += (double)(d * j4);
                s_den += (double)(j4 * j4);
            }
            a.J4[gi] = d;
            return v;
        }
    }
    return 0.f;
}

template <int E>
__global__ void __launch_bounds__(kMaxThreads) k_xpass(XArgs a, FftPlan pl) {
    extern __shared__ float2 lds[];
    const int t = threadIdx.x;
    const int P = pl.P;
    const int rr = t / P;
    const int p = t - rr * P;
    const int64_t row = (int64_t)blockIdx.x * a.RT + rr;
    const bool valid = (rr < a.RT) && (row < a.nrows);
    const int lbase = rr * a.lpad;
    const int L = pl.L;
    const bool in_spec = (a.mode == XM_C2R || a.mode == XM_RATIO || a.mode == XM_OMW_UPDATE ||
                          a.mode == XM_BIGGS_UPDATE);
    const bool out_spec = (a.mode == XM_R2C || a.mode == XM_PREP_U16 || a.mode == XM_PREP_F32 ||
                           a.mode == XM_RATIO || a.mode == XM_BIGGS_Y ||
                           ((a.mode == XM_OMW_UPDATE || a.mode == XM_BIGGS_UPDATE) && a.r2c_out));
    const int64_t srow = row * a.h1;      // spectrum row start
    const int64_t xrow = row * a.n1;      // real row start
    float2 v[E];
    double s_num = 0.0, s_den = 0.0;

    // ---- inverse transform along the row (spectrum -> real)
    if (in_spec) {
        if (a.packed) {
            const int M = L;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int m = p + e * P;
                float2 z = make_float2(0.f, 0.f);
                if (valid) {
                    const float2 xk = a.spec[srow + m];
                    const float2 xmk = cconj(a.spec[srow + (M - m)]);
                    const float2 ev = cadd(xk, xmk);
                    const float2 od = cmulc(csub(xk, xmk), a.rtw[m]);  // * W^{-m}
                    z = make_float2(ev.x - od.y, ev.y + od.x);          // ev + i od
                }
                v[e] = z;
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = p + e * P;
                float2 z = make_float2(0.f, 0.f);
                if (valid) z = (k < a.h1) ? a.spec[srow + k] : cconj(a.spec[srow + (L - k)]);
                v[e] = z;
            }
        }
        fft_run<E, true>(v, pl, p, lds, lbase, 1);
    }

    // ---- elementwise in real space
    if (a.packed) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = p + e * P;
            float2 o = make_float2(0.f, 0.f);
            if (valid) {
                const int64_t gi = xrow + 2 * m;
                o.x = x_elem(a, gi, v[e].x, s_num, s_den);
                o.y = x_elem(a, gi + 1, v[e].y, s_num, s_den);
            }
            v[e] = o;
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = p + e * P;
            float2 o = make_float2(0.f, 0.f);
            if (valid) o.x = x_elem(a, xrow + m, v[e].x, s_num, s_den);
            v[e] = o;
        }
    }

    // ---- forward transform along the row (real -> half spectrum)
    if (out_spec) {
        fft_run<E, false>(v, pl, p, lds, lbase, 1);
        if (a.packed) {
            const int M = L;
            // Z at home positions -> LDS, then X[k] = 1/2[(Z_k + Z*_{M-k}) - i W^k (Z_k - Z*_{M-k})]
#pragma unroll
            for (int e = 0; e < E; ++e) lds[lbase + p + e * P] = v[e];
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = p + e * P;
                const float2 A = lds[lbase + k];
                const float2 B = cconj(lds[lbase + (k == 0 ? 0 : M - k)]);
                const float2 s = cadd(A, B);
                const float2 d = cmul(csub(A, B), a.rtw[k]);
                // -i * d = (d.y, -d.x)
                const float2 X = make_float2(0.5f * (s.x + d.y), 0.5f * (s.y - d.x));
                if (valid) a.spec[srow + k] = X;
            }
            if (valid && p == 0) {
                const float2 z0 = lds[lbase];
                a.spec[srow + M] = make_float2(z0.x - z0.y, 0.f);
            }
            __syncthreads();
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = p + e * P;
                if (valid && k < a.h1) a.spec[srow + k] = v[e];
            }
        }
    }

    // ---- block partial sums for the Biggs step length (fixed order)
    if (a.mode == XM_BIGGS_UPDATE && a.acc) {
        __syncthreads();
        double* red = reinterpret_cast<double*>(lds);
        red[2 * t] = s_num;
        red[2 * t + 1] = s_den;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w && t + w < (int)blockDim.x) {
                red[2 * t] += red[2 * (t + w)];
                red[2 * t + 1] += red[2 * (t + w) + 1];
            }
            __syncthreads();
        }
        if (t == 0) {
            a.partials[2 * blockIdx.x] = red[0];
            a.partials[2 * blockIdx.x + 1] = red[1];
        }
    }
}

// ------------------------------------------------------------------ strided axes
enum CMode : int { CM_FWD = 0, CM_INV = 1, CM_CONV = 2, CM_CONVJ = 3 };

struct ColArgs {
    float2* data;
    const float2* otf;   // same layout as data (CM_CONV / CM_CONVJ)
    int64_t acount;      // columns along grid.x (contiguous index)
    int64_t bstride;     // offset between grid.y slices
    int64_t stride;      // element stride along the transform
    int T;               // columns per block
    int mode;
    float scale;         // applied with the OTF product
};

template <int E>
__global__ void __launch_bounds__(kMaxThreads) k_colpass(ColArgs a, FftPlan pl) {
    extern __shared__ float2 lds[];
    const int t = threadIdx.x;
    const int T = a.T;
    const int p = t / T;
    const int c = t - p * T;
    const int P = pl.P;
    const int64_t col = (int64_t)blockIdx.x * T + c;
    const bool valid = (p < P) && (col < a.acount);
    // per-lane base pointer plus a block-uniform step: addresses are
    // recomputed from (ptr0, e*step) instead of held live between load and store
    const int64_t step = (int64_t)P * a.stride;
    const int64_t off0 = col + (int64_t)blockIdx.y * a.bstride + (int64_t)p * a.stride;
    float2* __restrict__ ptr0 = a.data + (valid ? off0 : 0);
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = valid ? ptr0[e * step] : make_float2(0.f, 0.f);
    if (a.mode == CM_INV) {
        fft_run<E, true>(v, pl, p, lds, c, T);
    } else {
        fft_run<E, false>(v, pl, p, lds, c, T);
        if (a.mode >= CM_CONV) {
            const float2* __restrict__ h0 = a.otf + (valid ? off0 : 0);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                float2 h = valid ? h0[e * step] : make_float2(0.f, 0.f);
                h = make_float2(h.x * a.scale, h.y * a.scale);
                v[e] = (a.mode == CM_CONVJ) ? cmulc(v[e], h) : cmul(v[e], h);
            }
            fft_run<E, true>(v, pl, p, lds, c, T);
        }
    }
    if (valid) {
#pragma unroll
        for (int e = 0; e < E; ++e) ptr0[e * step] = v[e];
    }
}

}  // namespace pk
