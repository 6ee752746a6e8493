// pk_dft.hpp -- compile-time, register-resident small DFTs for gfx950.
//
// dft<R, INV>(x) transforms R complex values held in registers, natural order
// in and out.  Forward uses W = exp(-2*pi*i/R), inverse exp(+2*pi*i/R), no
// scaling.  Base butterflies are radix 2, 3, 4, 5; composite R = A*B is split
// Cooley-Tukey style (B-point DFTs over the A decimated sub-sequences, a
// twiddle by W_R^(n1*k2), then A-point DFTs), with every twiddle a constexpr
// constant so the compiler sees only adds and multiplies by literals.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

// Every device kernel includes this header first.  FP contraction is off for
// the rest of the translation unit and fused multiply-adds are written out
// (__builtin_fmaf), so the rounding of every operation is fixed by the source:
// kernels of different shape (compile-time and runtime plans, single and
// batched launches) that run the same operations produce identical bits.
#pragma clang fp contract(off)

namespace pk {

// ---------------------------------------------------------------- constexpr trig
// cos / sin of 2*pi*n/d, exact at multiples of pi/4, Taylor series elsewhere
// (argument reduced to [-pi, pi]; absolute error ~1e-15, far below FP32 ulp).
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
    double t = x, s = x;
    for (int k = 1; k < 24; ++k) {
        t *= -x * x / double((2 * k) * (2 * k + 1));
        s += t;
    }
    return s;
}
constexpr double taylor_cos(double x) {
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 24; ++k) {
        t *= -x * x / double((2 * k - 1) * (2 * k));
        s += t;
    }
    return s;
}
constexpr double cos2pi(long n, long d) {
    n %= d;
    if (n < 0) n += d;
    if (8 * n == 0) return 1.0;
    if (4 * n == d) return 0.0;
    if (2 * n == d) return -1.0;
    if (4 * n == 3 * d) return 0.0;
    if (2 * n > d) n -= d;
    return taylor_cos(2.0 * kPi * double(n) / double(d));
}
constexpr double sin2pi(long n, long d) {
    n %= d;
    if (n < 0) n += d;
    if (n == 0) return 0.0;
    if (4 * n == d) return 1.0;
    if (2 * n == d) return 0.0;
    if (4 * n == 3 * d) return -1.0;
    if (2 * n > d) n -= d;
    return taylor_sin(2.0 * kPi * double(n) / double(d));
}

// ---------------------------------------------------------------- complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(__builtin_fmaf(a.x, b.x, a.y * b.y), __builtin_fmaf(a.y, b.x, -(a.x * b.y)));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward quarter turn) or +i
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }
__device__ __forceinline__ float2 mul_pi(float2 a) { return make_float2(-a.y, a.x); }

template <int I, int N, class Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<I + 1, N>(fn);
    }
}

constexpr int first_factor(int r) {
    if (r % 4 == 0 && r > 4) return 4;
    for (int p = 2; p <= r; ++p)
        if (r % p == 0) return p;
    return r;
}

// Multiply a by W_R^e with the sign of the transform (compile-time e, R).
template <int E_, int R, bool INV>
__device__ __forceinline__ float2 twiddle_const(float2 a) {
    constexpr int e = ((E_ % R) + R) % R;
    if constexpr (e == 0) {
        return a;
    } else if constexpr (4 * e == R) {
        return INV ? mul_pi(a) : mul_mi(a);
    } else if constexpr (2 * e == R) {
        return make_float2(-a.x, -a.y);
    } else if constexpr (4 * e == 3 * R) {
        return INV ? mul_mi(a) : mul_pi(a);
    } else {
        constexpr float c = (float)cos2pi(e, R);
        constexpr float s = (float)(INV ? sin2pi(e, R) : -sin2pi(e, R));
        return make_float2(__builtin_fmaf(a.x, c, -(a.y * s)), __builtin_fmaf(a.x, s, a.y * c));
    }
}

// ---------------------------------------------------------------- base butterflies
template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
    const float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft3(float2& x0, float2& x1, float2& x2) {
    constexpr float s = 0.866025403784438646763723170752936183f;  // sin(2 pi / 3)
    const float2 t1 = cadd(x1, x2);
    const float2 d = csub(x1, x2);
    const float2 t2 = make_float2(__builtin_fmaf(-0.5f, t1.x, x0.x), __builtin_fmaf(-0.5f, t1.y, x0.y));
    x0 = cadd(x0, t1);
    // forward: X1 = t2 - i s d, X2 = t2 + i s d
    const float2 sd = make_float2(s * d.y, -s * d.x);  // -i s d
    if (INV) {
        x1 = csub(t2, sd);
        x2 = cadd(t2, sd);
    } else {
        x1 = cadd(t2, sd);
        x2 = csub(t2, sd);
    }
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
    const float2 s0 = cadd(x0, x2), d0 = csub(x0, x2);
    const float2 s1 = cadd(x1, x3), d1 = csub(x1, x3);
    const float2 r1 = INV ? mul_pi(d1) : mul_mi(d1);
    x0 = cadd(s0, s1);
    x2 = csub(s0, s1);
    x1 = cadd(d0, r1);
    x3 = csub(d0, r1);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2& x0, float2& x1, float2& x2, float2& x3, float2& x4) {
    constexpr float c1 = 0.309016994374947424102293417182819059f;   // cos(2 pi/5)
    constexpr float c2 = -0.809016994374947424102293417182819059f;  // cos(4 pi/5)
    constexpr float s1 = 0.951056516295153572116439333379382143f;   // sin(2 pi/5)
    constexpr float s2 = 0.587785252292473129168705954639072769f;   // sin(4 pi/5)
    const float2 a1 = cadd(x1, x4), b1 = csub(x1, x4);
    const float2 a2 = cadd(x2, x3), b2 = csub(x2, x3);
    const float2 t1 = make_float2(__builtin_fmaf(c2, a2.x, __builtin_fmaf(c1, a1.x, x0.x)),
                                  __builtin_fmaf(c2, a2.y, __builtin_fmaf(c1, a1.y, x0.y)));
    const float2 t2 = make_float2(__builtin_fmaf(c1, a2.x, __builtin_fmaf(c2, a1.x, x0.x)),
                                  __builtin_fmaf(c1, a2.y, __builtin_fmaf(c2, a1.y, x0.y)));
    const float2 u1 = make_float2(__builtin_fmaf(s2, b2.x, s1 * b1.x), __builtin_fmaf(s2, b2.y, s1 * b1.y));
    const float2 u2 = make_float2(__builtin_fmaf(-s1, b2.x, s2 * b1.x), __builtin_fmaf(-s1, b2.y, s2 * b1.y));
    x0 = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    // forward: X1 = t1 - i u1, X4 = t1 + i u1, X2 = t2 - i u2, X3 = t2 + i u2
    const float2 iu1 = mul_pi(u1), iu2 = mul_pi(u2);
    if (INV) {
        x1 = cadd(t1, iu1);
        x4 = csub(t1, iu1);
        x2 = cadd(t2, iu2);
        x3 = csub(t2, iu2);
    } else {
        x1 = csub(t1, iu1);
        x4 = cadd(t1, iu1);
        x2 = csub(t2, iu2);
        x3 = cadd(t2, iu2);
    }
}

// ---------------------------------------------------------------- generic DFT<R>
template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&x)[R]);

template <int R, bool INV>
struct DftImpl {
    static __device__ __forceinline__ void run(float2 (&x)[R]) {
        constexpr int A = first_factor(R);
        constexpr int B = R / A;
        float2 y[R];
        // B-point DFTs over n2 of x[n1 + A*n2], twiddled by W_R^(n1*k2)
        static_for<0, A>([&](auto n1c) {
            constexpr int n1 = decltype(n1c)::value;
            float2 t[B];
            static_for<0, B>([&](auto n2c) {
                constexpr int n2 = decltype(n2c)::value;
                t[n2] = x[n1 + A * n2];
            });
            dft<B, INV>(t);
            static_for<0, B>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                y[n1 * B + k2] = twiddle_const<n1 * k2, R, INV>(t[k2]);
            });
        });
        // A-point DFTs over n1 -> X[B*k1 + k2]
        static_for<0, B>([&](auto k2c) {
            constexpr int k2 = decltype(k2c)::value;
            float2 t[A];
            static_for<0, A>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                t[n1] = y[n1 * B + k2];
            });
            dft<A, INV>(t);
            static_for<0, A>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                x[B * k1 + k2] = t[k1];
            });
        });
    }
};

template <bool INV>
struct DftImpl<1, INV> {
    static __device__ __forceinline__ void run(float2 (&)[1]) {}
};
template <bool INV>
struct DftImpl<2, INV> {
    static __device__ __forceinline__ void run(float2 (&x)[2]) { dft2<INV>(x[0], x[1]); }
};
template <bool INV>
struct DftImpl<3, INV> {
    static __device__ __forceinline__ void run(float2 (&x)[3]) { dft3<INV>(x[0], x[1], x[2]); }
};
template <bool INV>
struct DftImpl<4, INV> {
    static __device__ __forceinline__ void run(float2 (&x)[4]) { dft4<INV>(x[0], x[1], x[2], x[3]); }
};
template <bool INV>
struct DftImpl<5, INV> {
    static __device__ __forceinline__ void run(float2 (&x)[5]) { dft5<INV>(x[0], x[1], x[2], x[3], x[4]); }
};

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&x)[R]) {
    DftImpl<R, INV>::run(x);
}

}  // namespace pk
