// pk_fft_dev.hpp -- device side of the gfx950 deconvolution engine.
//
// Volumes are MATLAB column-major: dim 1 (n1) contiguous.  Real volumes hold
// N = n1*n2*n3 floats; spectra hold the half spectrum along dim 1,
// h1 = floor(n1/2)+1 complex values per row stored at a row pitch hp (h1
// rounded up to 16 float2 = 128 bytes, so every row and every 16-column tile
// of the strided passes starts on a cache-line boundary); the pad columns
// are never read or written.
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
#include "pk_misc.hpp"

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

// LDS slot of position pos.  Column tiles (ROWS=false) interleave T columns:
// base + pos*T.  Row tiles (ROWS=true) keep a row contiguous with one pad
// slot after every 16 positions, so the stride-R writes of the first stage
// (lane stride R*8 bytes) and the home-position reads spread over all banks.
template <bool ROWS>
__device__ __forceinline__ int lds_slot(int lbase, int pos, int lstride) {
    if constexpr (ROWS) return lbase + pos + (pos >> 4);
    else return lbase + pos * lstride;
}
__host__ __device__ constexpr int row_pitch(int L) { return L + (L >> 4) + 1; }

// Radix of the next Stockham stage: the first entry of this list that divides
// both the remaining length and E (the host planner and the compile-time row
// kernels factor with the same rule, so both run the same stage sequence).
constexpr int kRadixOrder[] = {32, 30, 27, 25, 24, 20, 18, 16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2};
__host__ __device__ constexpr int radix_for(int rem, int E) {
    for (int r : kRadixOrder)
        if (E % r == 0 && rem % r == 0) return r;
    return 0;
}

// Inter-stage twiddles of one radix-R butterfly: u[m] *= W_L^(m k tws).
// R <= 4: each power read from the length-L table (exact table values);
// larger R: one table read, the rest by binary powering (w_2m = w_m^2,
// w_2m+1 = w_2m w_1; chain depth <= log2 R).
template <int R, bool INV>
__device__ __forceinline__ void stage_twiddle(float2 (&u)[R], const float2* __restrict__ tw, int k, int tws) {
    if constexpr (R <= 4) {
        static_for<1, R>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const float2 w = tw[m * k * tws];
            u[m] = INV ? cmulc(u[m], w) : cmul(u[m], w);
        });
    } else {
        float2 w[R];
        w[1] = tw[k * tws];
        static_for<2, R>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            w[m] = (m % 2 == 0) ? cmul(w[m / 2], w[m / 2]) : cmul(w[m - 1], w[1]);
        });
        static_for<1, R>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            u[m] = INV ? cmulc(u[m], w[m]) : cmul(u[m], w[m]);
        });
    }
}

// ------------------------------------------------------------------ one stage
template <int E, int R, bool INV, bool ROWS>
__device__ __forceinline__ void stage_r(float2 (&v)[E], const FftPlan& pl, int s, int p, float2* lds, int lbase,
                                        int lstride, bool last, bool act) {
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
        if (s != 0) stage_twiddle<R, INV>(u, pl.tw, k, tws);
        dft<R, INV>(u);
        if (last) {
            static_for<0, R>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                v[i + q * Q] = u[q];
            });
        } else if (act) {
            const int ob = g * Ns * R + k;
            static_for<0, R>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                lds[lds_slot<ROWS>(lbase, ob + q * Ns, lstride)] = u[q];
            });
        }
    }
}

// A stage whose radix R does not divide E (plans with such stages keep few
// values per thread for lengths with the primes 2, 3 and 5): the L / R
// butterflies are dealt round-robin over the P threads, at most ceil(E / R)
// rounds each (the last predicated), and read their inputs from LDS, where
// the previous stage left its outputs (stage 0 first stores the home
// registers).  Outputs go to the usual Stockham positions in LDS.  Each
// butterfly does the same operations as in stage_r.
template <int E, int R, bool INV, bool ROWS>
__device__ __forceinline__ void stage_mm(float2 (&v)[E], const FftPlan& pl, int s, int p, float2* lds, int lbase,
                                         int lstride, bool act) {
    constexpr int NB = (E + R - 1) / R;
    const int P = pl.P;
    const int NBF = pl.L / R;
    const int Ns = pl.Ns[s];
    const int tws = pl.tws[s];
    const uint32_t mag = pl.mag[s];
    if (s == 0) {
        if (act) {
#pragma unroll
            for (int e = 0; e < E; ++e) lds[lds_slot<ROWS>(lbase, p + e * P, lstride)] = v[e];
        }
        __syncthreads();
    }
    float2 u[NB][R];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = p + b * P;
        const int jj = j < NBF ? j : 0;  // idle round: an in-bounds read
        static_for<0, R>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            u[b][m] = act ? lds[lds_slot<ROWS>(lbase, jj + m * NBF, lstride)] : make_float2(0.f, 0.f);
        });
    }
    __syncthreads();  // every input read before any output lands
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = p + b * P;
        if (act && j < NBF) {
            int g, k;
            if (s == 0) {
                g = j;
                k = 0;
            } else {
                g = (int)__umulhi((uint32_t)j, mag);
                k = j - g * Ns;
            }
            if (s != 0) stage_twiddle<R, INV>(u[b], pl.tw, k, tws);
            dft<R, INV>(u[b]);
            const int ob = g * Ns * R + k;
            static_for<0, R>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                lds[lds_slot<ROWS>(lbase, ob + q * Ns, lstride)] = u[b][q];
            });
        }
    }
}

// Radices a plan may run from LDS (stage_mm), largest first.
constexpr int kMixedRadices[] = {5, 4, 3, 2};
// A plan with at least one stage whose radix does not divide E.
__host__ __device__ inline bool plan_mixed_stages(const FftPlan& pl) {
    for (int s = 0; s < pl.nst; ++s)
        if (pl.E % pl.R[s] != 0) return true;
    return false;
}

#define PK_RADIX_CASE(RR)                                                                                    \
    case RR:                                                                                                 \
        if constexpr (E % RR == 0) stage_r<E, RR, INV, ROWS>(v, pl, s, p, lds, lbase, lstride, last, act);   \
        else if constexpr (MIXED && RR <= 5) stage_mm<E, RR, INV, ROWS>(v, pl, s, p, lds, lbase, lstride, act); \
        break;

// Full transform of the E home values of every thread (all threads of the
// block must call it: it contains barriers).  act = false marks a thread that
// owns no tile slot (batched launches with a block larger than the chunk's
// tile): it joins the barriers but never touches LDS.  MIXED: the plan may
// hold stages whose radix (2..5) does not divide E.
template <int E, bool INV, bool ROWS, bool MIXED = false>
__device__ __forceinline__ void fft_run(float2 (&v)[E], const FftPlan& pl, int p, float2* lds, int lbase,
                                        int lstride, bool act = true) {
    for (int s = 0; s < pl.nst; ++s) {
        const bool last = (s == pl.nst - 1);
        const bool mm = MIXED && (E % pl.R[s] != 0);
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
        if (!last || mm) {  // outputs in LDS: back to home registers unless the next stage reads LDS itself
            __syncthreads();
            const bool next_mm = MIXED && !last && (E % pl.R[s + 1] != 0);
            if (!next_mm) {
#pragma unroll
                for (int e = 0; e < E; ++e)
                    if (act) v[e] = lds[lds_slot<ROWS>(lbase, p + e * pl.P, lstride)];
                __syncthreads();
            }
        }
    }
}
#undef PK_RADIX_CASE

// ------------------------------------------------------------------ compile-time rows
// The same Stockham stages with L and E fixed at compile time: radices,
// strides, twiddle strides and LDS offsets are constants.  When a row's P
// threads lie inside one wavefront (P <= 64) the exchanges only need the
// wavefront's own ordering: LDS operations of one wavefront execute in
// order, so a wavefront-scope fence (a compiler ordering point) replaces the
// workgroup barrier.
//
// A compile-time stage may also take a radix that does not divide E (E then
// need not hold every prime of L, so long rows with the primes 2, 3 and 5
// keep few values per thread).  Such a stage reads its butterflies' inputs
// from LDS instead of home registers: the L / R butterflies are dealt
// round-robin over the P threads (the last round predicated when P does not
// divide L / R), and the outputs go to the usual Stockham positions.  Each
// butterfly does the same operations either way, so the bits of a transform
// depend only on its radix sequence.
__host__ __device__ constexpr int ct_radix_for(int rem, int E) {
    const int r = radix_for(rem, E);
    if (r) return r;
    for (int q : kRadixOrder)
        if (rem % q == 0) return q;
    return 0;
}
__host__ __device__ constexpr int ct_nst(int L, int E) {
    int n = 0;
    for (int rem = L; rem > 1; ++n) rem /= ct_radix_for(rem, E);
    return n;
}
__host__ __device__ constexpr int ct_radix(int L, int E, int s) {
    int rem = L;
    for (int i = 0; i < s; ++i) rem /= ct_radix_for(rem, E);
    return ct_radix_for(rem, E);
}
__host__ __device__ constexpr int ct_ns(int L, int E, int s) {
    int ns = 1, rem = L;
    for (int i = 0; i < s; ++i) {
        const int r = ct_radix_for(rem, E);
        ns *= r;
        rem /= r;
    }
    return ns;
}
template <int P>
__device__ __forceinline__ void row_sync() {
    if constexpr (P <= 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// One stage.  Data enter in home registers, except after a stage whose
// successor reads LDS itself (the outputs are then left in LDS); they leave
// in home registers after the last stage.
template <int L, int E, int S, bool INV>
__device__ __forceinline__ void ct_stage(float2 (&v)[E], int p, float2* lds, int lbase, const float2* __restrict__ tw) {
    constexpr int R = ct_radix(L, E, S);
    constexpr int Ns = ct_ns(L, E, S);
    constexpr int P = L / E, tws = L / (Ns * R);
    constexpr bool last = (S == ct_nst(L, E) - 1);
    constexpr bool matched = (R > 0 && E % R == 0);
    constexpr bool next_from_lds = !last && (E % ct_radix(L, E, S + 1) != 0);
    static_assert(R > 0, "no stage radix");
    if constexpr (matched) {
        constexpr int Q = E / R;
        static_for<0, Q>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const int j = p + i * P;
            const int g = j / Ns;
            const int k = j - g * Ns;
            float2 u[R];
            static_for<0, R>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                u[m] = v[i + m * Q];
            });
            if constexpr (S != 0) stage_twiddle<R, INV>(u, tw, k, tws);
            dft<R, INV>(u);
            if constexpr (last) {
                static_for<0, R>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    v[i + q * Q] = u[q];
                });
            } else {
                const int ob = g * Ns * R + k;
                static_for<0, R>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    lds[lds_slot<true>(lbase, ob + q * Ns, 1)] = u[q];
                });
            }
        });
    } else {
        constexpr int NBF = L / R;             // butterflies of the stage
        constexpr int NB = (NBF + P - 1) / P;  // rounds per thread
        if constexpr (S == 0) {  // inputs still in home registers
            static_for<0, E>([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                lds[lds_slot<true>(lbase, p + e * P, 1)] = v[e];
            });
            row_sync<P>();
        }
        float2 u[NB][R];
        static_for<0, NB>([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            const int j = p + b * P;
            const int jj = (NBF % P == 0 || b + 1 < NB || j < NBF) ? j : 0;  // idle slot: an in-bounds read
            static_for<0, R>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                u[b][m] = lds[lds_slot<true>(lbase, jj + m * NBF, 1)];
            });
        });
        row_sync<P>();  // every input read before any output lands
        static_for<0, NB>([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            const int j = p + b * P;
            if (NBF % P == 0 || b + 1 < NB || j < NBF) {
                const int g = j / Ns;
                const int k = j - g * Ns;
                if constexpr (S != 0) stage_twiddle<R, INV>(u[b], tw, k, tws);
                dft<R, INV>(u[b]);
                const int ob = g * Ns * R + k;
                static_for<0, R>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    lds[lds_slot<true>(lbase, ob + q * Ns, 1)] = u[b][q];
                });
            }
        });
    }
    if constexpr (!(matched && last)) {  // outputs in LDS
        row_sync<P>();
        if constexpr (!next_from_lds) {
            static_for<0, E>([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                v[e] = lds[lds_slot<true>(lbase, p + e * P, 1)];
            });
            row_sync<P>();
        }
    }
}

template <int L, int E, bool INV, int S = 0>
__device__ __forceinline__ void ct_fft(float2 (&v)[E], int p, float2* lds, int lbase, const float2* __restrict__ tw) {
    if constexpr (S < ct_nst(L, E)) {
        ct_stage<L, E, S, INV>(v, p, lds, lbase, tw);
        ct_fft<L, E, INV, S + 1>(v, p, lds, lbase, tw);
    }
}

// ------------------------------------------------------------------ x-axis (rows)
enum XMode : int {
    XM_R2C = 0,          // real src -> spectrum
    XM_C2R = 1,          // spectrum -> real dst (unnormalised)
    XM_PREP_U16 = 2,     // raw uint16: I = sat(raw - bg16); I -> dst_I, J; spectrum of I
    XM_PREP_F32 = 3,     // raw float: I = max(raw - bg, 0); I -> dst_I, J; spectrum of I
    XM_RATIO = 4,        // spectrum -> c; r = I / max(c, eps); spectrum of r
    XM_OMW_UPDATE = 5,   // spectrum -> t; J = max(t * J, 0); [spectrum of J]
    XM_BIGGS_Y = 6,      // Y = max(J2 + lambda (J2 - J3), 0) -> Y; spectrum of Y
    XM_BIGGS_UPDATE = 7, // spectrum -> t; Jn = max(Y * t, 0) -> Jn; J4 = Jn - Y; dot partials
    // last update with the epilogue fused (whole volume, single volume only):
    // the new estimate goes through the epilogue into ep.out16 / ep.outf
    XM_OMW_FINAL = 8,    // mode 5 without the estimate store or the forward transform
    XM_BIGGS_FINAL = 9,  // mode 7 without Jn, J4, the dot partials or the forward transform
    // uint16 input (compile-time row kernels only): I = sat(raw - bg16) holds
    // integers below 2^16, kept as uint16 (I16), half the bytes of float I
    XM_RATIO_U16 = 10,       // mode 4 reading I from I16
    XM_OMW_UPDATE_U16 = 11,  // mode 5 reading the estimate from I16 (the first update: J_1 = I)
};
// the iteration mode a fused final or uint16-operand mode extends
constexpr int xm_base(int m) {
    return m == XM_OMW_FINAL || m == XM_OMW_UPDATE_U16 ? XM_OMW_UPDATE
           : m == XM_BIGGS_FINAL                      ? XM_BIGGS_UPDATE
           : m == XM_RATIO_U16                        ? XM_RATIO
                                                      : m;
}
constexpr bool xm_final(int m) { return m == XM_OMW_FINAL || m == XM_BIGGS_FINAL; }
constexpr bool xm_u16(int m) { return m == XM_RATIO_U16 || m == XM_OMW_UPDATE_U16; }

struct XArgs {
    int n1;            // real row length
    int h1;            // n1/2 + 1
    int hp;            // spectrum row pitch (float2)
    int packed;        // n1 even: transform length n1/2 over (x[2m], x[2m+1])
    int RT;            // rows per block
    int lpad;          // LDS row pitch (float2)
    int mode;
    int r2c_out;       // modes 5/7: also produce the spectrum of the new estimate
    int acc;           // mode 7: accumulate the two dot products
    int write_j4;      // mode 7: store J4 = Jn - Y
    int64_t nrows;     // n2 * n3
    float bg;          // background (float) / bg16 for uint16
    float eps;         // MATLAB eps as single
    float2* spec;      // half spectrum (in and/or out)
    const float2* rtw; // W_{n1}^k, k in [0, n1/2]  (packed post-processing)
    const float* src;  // XM_R2C input
    float* dst;        // XM_C2R output
    const uint16_t* raw16;
    const float* rawf;
    float* I;          // background-subtracted input
    uint16_t* I16;     // uint16 input: I as uint16, written by the prep instead of I when set
    float* J;          // estimate (OMW J, Biggs J2 for mode 6); modes 2/3 store it only when non-null
    const float* Jin;  // mode 5: read the estimate from here instead of J when non-null (first iteration: I)
    const float* J3;   // Biggs x_{k-2}
    float* Y;          // Biggs prediction
    float* Jn;         // Biggs new estimate
    float* J4;         // Biggs x_{k-1} - y_{k-1}
    const float* lambda;  // device scalar
    double* partials;     // 2 doubles per block
    EpiArgs ep;           // modes 8/9: epilogue operands and the (cropped) output
    // modes 2/3: raw16 / rawf in the caller's unpadded s0 x s1 x s2 layout
    // inside the n1 x pn2 x n3 plan (zeros past its edge); s0 = 0: same layout
    int64_t s0, s1, s2, pn2;
};

// Cache policy of the passes' streaming accesses (compile-time).  A
// nontemporal access (the `nt` bit) tells the caches not to keep the line for
// reuse: every pass touches each value once, and the next pass reads it back
// only after the whole volume has gone through.
//   PK_NT_COL: bit 0 strided-pass spectrum loads, bit 1 their stores, bit 2 OTF loads
//   PK_NT_ROW: bit 0 row-pass spectrum loads, bit 1 their stores, bit 2 real
//              operand loads, bit 3 real stores
// Default (PK_NT_COL 6): nontemporal stores in the strided passes (tiles of
// 16 or more columns: whole 128-byte row segments) and nontemporal OTF loads
// (each OTF is read once per iteration and is as large as the spectrum, which
// the next pass reads back).  Measured against all accesses temporal (ms per
// call / Gvoxel/s): 512x512x256 OMW-2 2.578 -> 2.546 with the stores, -> 2.488
// with the OTF loads too (z_conv 0.168 -> 0.153 ms); 64 chunks OMW-2 18.33 ->
// 18.71 / Biggs-10 3.39 -> 3.52 with the stores (the batched kernels keep
// plain OTF loads: 0.5-0.7 % lower with nontemporal ones, interleaved runs);
// 2048x2048x512 23.45 -> 23.72 -> 23.84.  Not
// kept: spectrum loads as well (2.587 ms), row-pass operand loads (2.518),
// stores in every pass (2.584), every access (2.583).
#ifndef PK_NT_COL
#define PK_NT_COL 6
#endif
#ifndef PK_NT_ROW
#define PK_NT_ROW 0
#endif
typedef float pk_vf2 __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ float2 ld_f2(const float2* p) {
    if constexpr (NT) {
        const pk_vf2 v = __builtin_nontemporal_load(reinterpret_cast<const pk_vf2*>(p));
        return make_float2(v.x, v.y);
    } else {
        return *p;
    }
}
template <bool NT>
__device__ __forceinline__ void st_f2(float2* p, float2 v) {
    if constexpr (NT) {
        pk_vf2 w;
        w.x = v.x;
        w.y = v.y;
        __builtin_nontemporal_store(w, reinterpret_cast<pk_vf2*>(p));
    } else {
        *p = v;
    }
}
template <bool NT>
__device__ __forceinline__ float ld_f1(const float* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void st_f1(float* p, float v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}
constexpr bool kNtColLd = (PK_NT_COL & 1) != 0, kNtColSt = (PK_NT_COL & 2) != 0, kNtOtfLd = (PK_NT_COL & 4) != 0;
constexpr bool kNtRowLd = (PK_NT_ROW & 1) != 0, kNtRowSt = (PK_NT_ROW & 2) != 0;
constexpr bool kNtRealLd = (PK_NT_ROW & 4) != 0, kNtRealSt = (PK_NT_ROW & 8) != 0;

__device__ __forceinline__ float fmax0(float a) { return a > 0.f ? a : 0.f; }  // max(a,0), NaN -> 0
__device__ __forceinline__ float2 fmax0(float2 a) { return make_float2(fmax0(a.x), fmax0(a.y)); }

// Real-space access of element e of a row: packed rows pair (x[2m], x[2m+1])
// into one float2 (8-byte vector access); odd rows carry x[m] in .x.
__device__ __forceinline__ float2 ld_real(const float* __restrict__ base, int64_t xrow, int m, bool packed) {
    if (packed) return reinterpret_cast<const float2*>(base)[(xrow >> 1) + m];
    return make_float2(base[xrow + m], 0.f);
}
__device__ __forceinline__ void st_real(float* __restrict__ base, int64_t xrow, int m, bool packed, float2 v) {
    if (packed) reinterpret_cast<float2*>(base)[(xrow >> 1) + m] = v;
    else base[xrow + m] = v.x;
}
__device__ __forceinline__ void st_u16(uint16_t* __restrict__ base, int64_t xrow, int m, bool packed, float2 v) {
    if (packed) reinterpret_cast<uint32_t*>(base)[(xrow >> 1) + m] = (uint32_t)v.x | ((uint32_t)v.y << 16);
    else base[xrow + m] = (uint16_t)v.x;
}
__device__ __forceinline__ float2 ld_u16(const uint16_t* __restrict__ base, int64_t xrow, int m, bool packed) {
    if (packed) {
        const uint32_t w = reinterpret_cast<const uint32_t*>(base)[(xrow >> 1) + m];
        return make_float2((float)(w & 0xffffu), (float)(w >> 16));
    }
    return make_float2((float)base[xrow + m], 0.f);
}

// Raw input of a padded volume read from the caller's unpadded layout: value
// pair at x0, x0 + 1 (one value when !pair) of plan row `row`; positions
// outside the s0 x s1 x s2 source read as 0 (the zero padding).
struct RawRow {
    int64_t src;  // source index of x = 0
    bool in;      // the row lies inside the source in dims 2 and 3
};
__device__ __forceinline__ RawRow raw_row(const XArgs& a, int64_t row, int64_t n2) {
    const int64_t y = row % n2, z = row / n2;
    return RawRow{(y + a.s1 * z) * a.s0, y < a.s1 && z < a.s2};
}
// A pair inside the source row whose first value sits on a 2-value boundary
// is read as one word (4 bytes for uint16, 8 for float); the values are the
// same either way.
template <class T>
__device__ __forceinline__ float2 ld_raw_padded(const T* __restrict__ base, const XArgs& a, const RawRow& r,
                                                int64_t x0, bool pair) {
    float2 v = make_float2(0.f, 0.f);
    if (!r.in) return v;
    const T* __restrict__ q = base + (r.src + x0);
    if (pair && x0 + 1 < a.s0) {
        if ((reinterpret_cast<uintptr_t>(q) & (2 * sizeof(T) - 1)) == 0) {
            if constexpr (sizeof(T) == 2) {
                const uint32_t w = *reinterpret_cast<const uint32_t*>(q);
                v.x = (float)(w & 0xffffu);
                v.y = (float)(w >> 16);
            } else {
                const float2 f = *reinterpret_cast<const float2*>(q);
                v.x = (float)f.x;
                v.y = (float)f.y;
            }
        } else {
            v.x = (float)q[0];
            v.y = (float)q[1];
        }
    } else if (x0 < a.s0) {
        v.x = (float)q[0];
    }
    return v;
}

// ---- per-element arithmetic shared by the generic and the compile-time row kernels
// (one definition, so both produce bit-identical results)

// a / b: reciprocal plus one Newton correction of the quotient
// (q = a r, e = a - b q, q += e r); within 1 ulp of the IEEE quotient for the
// operand range here (b >= eps, finite a).
__device__ __forceinline__ float div_nr(float a, float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = a * r;
    const float e = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(e, r, q);
}
// C2R pre-step of a packed row: Z_m = (X_m + X*_{M-m}) + i W^{-m} (X_m - X*_{M-m})
__device__ __forceinline__ float2 c2r_pre(float2 xk, float2 xmk_conj, float2 w) {
    const float2 ev = cadd(xk, xmk_conj);
    const float2 od = cmulc(csub(xk, xmk_conj), w);
    return make_float2(ev.x - od.y, ev.y + od.x);
}
// R2C post-step: X_k = 1/2 [(Z_k + Z*_{M-k}) - i W^k (Z_k - Z*_{M-k})]
__device__ __forceinline__ float2 r2c_post(float2 A, float2 Bc, float2 w) {
    const float2 s = cadd(A, Bc);
    const float2 d = cmul(csub(A, Bc), w);
    return make_float2(0.5f * (s.x + d.y), 0.5f * (s.y - d.x));
}
// Real-space step of one value pair (v: inverse-transform output, opA/opB:
// the mode's real-space operands); returns the forward-transform input.
template <int MODE>
__device__ __forceinline__ float2 x_elem(float2 v, float2 opA, float2 opB, float bg, float eps, float lam) {
    if constexpr (MODE == XM_R2C) return opA;
    if constexpr (MODE == XM_PREP_U16 || MODE == XM_PREP_F32)
        // uint16: I - uint16(Background) saturates at 0; single: max(I - Background, 0)
        return fmax0(make_float2(opA.x - bg, opA.y - bg));
    if constexpr (MODE == XM_RATIO)  // I ./ max(CX, eps)
        return make_float2(div_nr(opA.x, v.x > eps ? v.x : eps), div_nr(opA.y, v.y > eps ? v.y : eps));
    if constexpr (MODE == XM_OMW_UPDATE) return fmax0(make_float2(v.x * opA.x, v.y * opA.y));
    if constexpr (MODE == XM_BIGGS_Y)
        return fmax0(make_float2(opA.x + lam * (opA.x - opB.x), opA.y + lam * (opA.y - opB.y)));
    if constexpr (MODE == XM_BIGGS_UPDATE) return fmax0(make_float2(opA.x * v.x, opA.y * v.y));
    return v;
}

// One x pass, specialised on its mode: [C2R] -> elementwise -> [R2C].
template <int E, int MODE_>
__device__ __forceinline__ void xpass_body(const XArgs& a, const FftPlan& pl, float2* lds, int64_t blk) {
    constexpr int MODE = xm_base(MODE_);
    constexpr bool EPI = xm_final(MODE_);  // fused epilogue
    constexpr bool IN = (MODE == XM_C2R || MODE == XM_RATIO || MODE == XM_OMW_UPDATE || MODE == XM_BIGGS_UPDATE);
    constexpr bool OUT_ALWAYS = (MODE == XM_R2C || MODE == XM_PREP_U16 || MODE == XM_PREP_F32 ||
                                 MODE == XM_RATIO || MODE == XM_BIGGS_Y);
    constexpr bool OUT_MAYBE = (MODE == XM_OMW_UPDATE || MODE == XM_BIGGS_UPDATE);
    const int t = threadIdx.x;
    const int P = pl.P;
    const int rr = t / P;
    const int p = t - rr * P;
    const int64_t row = blk * a.RT + rr;
    const bool act = rr < a.RT;  // owns an LDS row slot
    const bool valid = act && (row < a.nrows);
    const int lbase = act ? rr * a.lpad : 0;
    const int L = pl.L;
    const bool packed = a.packed != 0;
    const int64_t srow = row * a.hp;  // spectrum row start
    const int64_t xrow = row * a.n1;  // real row start

    // ---- prefetch the real-space operands so their latency overlaps the C2R
    RawRow rr_src{0, false};
    if constexpr (MODE == XM_PREP_U16 || MODE == XM_PREP_F32) {
        if (a.s0 && valid) rr_src = raw_row(a, row, a.pn2);
    }
    float2 opA[E], opB[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = p + e * P;
        opA[e] = make_float2(0.f, 0.f);
        opB[e] = make_float2(0.f, 0.f);
        if (!valid) continue;
        if constexpr (MODE == XM_R2C) opA[e] = ld_real(a.src, xrow, m, packed);
        if constexpr (MODE == XM_PREP_U16)
            opA[e] = a.s0 ? ld_raw_padded(a.raw16, a, rr_src, packed ? 2 * (int64_t)m : m, packed)
                          : ld_u16(a.raw16, xrow, m, packed);
        if constexpr (MODE == XM_PREP_F32)
            opA[e] = a.s0 ? ld_raw_padded(a.rawf, a, rr_src, packed ? 2 * (int64_t)m : m, packed)
                          : ld_real(a.rawf, xrow, m, packed);
        if constexpr (MODE == XM_RATIO) opA[e] = ld_real(a.I, xrow, m, packed);
        if constexpr (MODE == XM_OMW_UPDATE) opA[e] = ld_real(a.Jin ? a.Jin : a.J, xrow, m, packed);
        if constexpr (MODE == XM_BIGGS_Y) {
            opA[e] = ld_real(a.J, xrow, m, packed);
            opB[e] = ld_real(a.J3, xrow, m, packed);
        }
        if constexpr (MODE == XM_BIGGS_UPDATE) {
            opA[e] = ld_real(a.Y, xrow, m, packed);
            if (!EPI && a.acc) opB[e] = ld_real(a.J4, xrow, m, packed);
        }
    }

    // ---- inverse transform along the row (half spectrum -> real)
    float2 v[E];
    if constexpr (IN) {
        if (packed) {
            const int M = L;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int m = p + e * P;
                float2 z = make_float2(0.f, 0.f);
                if (valid) z = c2r_pre(a.spec[srow + m], cconj(a.spec[srow + (M - m)]), a.rtw[m]);
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
        fft_run<E, true, true>(v, pl, p, lds, lbase, 1, act);
    }

    // ---- elementwise step in real space
    double s_num = 0.0, s_den = 0.0;
    float lam = 0.f;
    if constexpr (MODE == XM_BIGGS_Y) lam = *a.lambda;
    EpiRow er{0, 0, false};
    if constexpr (EPI) {
        if (valid) er = epi_row(a.ep, row);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = p + e * P;
        float2 o = make_float2(0.f, 0.f);
        if (valid) {
            if constexpr (MODE == XM_C2R) {
                st_real(a.dst, xrow, m, packed, v[e]);
            } else {
                o = x_elem<MODE>(v[e], opA[e], opB[e], a.bg, a.eps, lam);
            }
            if constexpr (MODE == XM_PREP_U16 || MODE == XM_PREP_F32) {
                if (MODE == XM_PREP_U16 && a.I16) st_u16(a.I16, xrow, m, packed, o);
                else st_real(a.I, xrow, m, packed, o);
                if (a.J) st_real(a.J, xrow, m, packed, o);
            } else if constexpr (MODE == XM_OMW_UPDATE) {
                if constexpr (EPI) {
                    if (er.keep) epi_put(a.ep, er, packed ? 2 * (int64_t)m : m, packed, o);
                } else {
                    st_real(a.J, xrow, m, packed, o);
                }
            } else if constexpr (MODE == XM_BIGGS_Y) {
                st_real(a.Y, xrow, m, packed, o);
            } else if constexpr (MODE == XM_BIGGS_UPDATE) {
                const float2 y = opA[e];
                if constexpr (EPI) {
                    if (er.keep) epi_put(a.ep, er, packed ? 2 * (int64_t)m : m, packed, o);
                } else {
                    st_real(a.Jn, xrow, m, packed, o);
                }
                const float2 d = make_float2(o.x - y.x, o.y - y.y);
                if (!EPI && a.acc) {
                    const float2 j4 = opB[e];
                    s_num += (double)(d.x * j4.x);
                    s_den += (double)(j4.x * j4.x);
                    if (packed) {
                        s_num += (double)(d.y * j4.y);
                        s_den += (double)(j4.y * j4.y);
                    }
                }
                if (!EPI && a.write_j4) st_real(a.J4, xrow, m, packed, d);
            }
            if (!packed) o.y = 0.f;  // odd rows: the forward transform input is real
        }
        v[e] = o;
    }

    // ---- forward transform along the row (real -> half spectrum)
    bool out_spec = OUT_ALWAYS;
    if constexpr (OUT_MAYBE) out_spec = !EPI && a.r2c_out != 0;
    if (out_spec) {
        fft_run<E, false, true>(v, pl, p, lds, lbase, 1, act);
        if (packed) {
            const int M = L;
            // Z at home positions -> LDS, then X[k] from Z_k and Z*_{M-k}
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (act) lds[lds_slot<true>(lbase, p + e * P, 1)] = v[e];
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = p + e * P;
                if (!act) continue;
                const float2 A = lds[lds_slot<true>(lbase, k, 1)];
                const float2 B = cconj(lds[lds_slot<true>(lbase, k == 0 ? 0 : M - k, 1)]);
                if (valid) a.spec[srow + k] = r2c_post(A, B, a.rtw[k]);
            }
            if (valid && p == 0) {
                const float2 z0 = lds[lbase];  // slot of position 0
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

    // ---- block partial sums of the Biggs step length (fixed order)
    if constexpr (MODE == XM_BIGGS_UPDATE && !EPI) {
        if (a.acc) {
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
                a.partials[2 * blk] = red[0];
                a.partials[2 * blk + 1] = red[1];
            }
        }
    }
}

// Compile-time row kernels load their real-space operands after the inverse
// transform when true (fewer live registers, exposed load latency).
constexpr bool kXrowLateOperands = false;

// Rows per 256-thread block of the compile-time row kernels.  A row of
// P <= 64 threads never straddles a wavefront: each wavefront holds
// floor(64 / P) rows, and the 64 mod P lanes left over (none when P divides
// 64) run a dummy row in an LDS slot of their wavefront's own and store
// nothing.  Longer rows (P = 128, 256) span whole wavefronts.
__host__ __device__ constexpr int xrow_tile(int P) { return P <= 64 ? 4 * (64 / P) : kMaxThreads / P; }
// LDS rows the kernel addresses: the row tile plus one dummy row per wavefront
// when lanes are left over.
__host__ __device__ constexpr int xrow_lds_rows(int P) {
    return xrow_tile(P) + ((P <= 64 && 64 % P != 0) ? kMaxThreads / 64 : 0);
}

// The same pass for an even row of n1 = 2L real values (ODD: an odd row of
// n1 = L real values, transformed at full length as complex values with zero
// imaginary parts, as the runtime-plan kernel does) with L and E fixed at
// compile time and 256-thread blocks of xrow_tile(P) rows.  Rows past the end
// of the volume (and the left-over lanes) read the last row (in bounds) and
// store nothing; every wave runs the same instruction stream.
template <int L, int E, int MODE_, bool ODD = false>
__device__ __forceinline__ void xrow_body(const XArgs& a, const float2* __restrict__ tw, float2* lds, int64_t blk) {
    constexpr int MODE = xm_base(MODE_);
    constexpr bool EPI = xm_final(MODE_);  // fused epilogue
    constexpr bool IN = (MODE == XM_C2R || MODE == XM_RATIO || MODE == XM_OMW_UPDATE || MODE == XM_BIGGS_UPDATE);
    constexpr bool OUT_ALWAYS = (MODE == XM_R2C || MODE == XM_PREP_U16 || MODE == XM_PREP_F32 ||
                                 MODE == XM_RATIO || MODE == XM_BIGGS_Y);
    constexpr bool OUT_MAYBE = (MODE == XM_OMW_UPDATE || MODE == XM_BIGGS_UPDATE);
    constexpr int P = L / E;
    constexpr int RT = xrow_tile(P);
    constexpr int LP = row_pitch(L);
    constexpr int H1 = ODD ? (L + 1) / 2 : L + 1;  // half-spectrum length of the row
    constexpr int HP = (H1 + 15) / 16 * 16;       // its pitch
    static_assert(P * E == L && (P <= 64 || RT * P == kMaxThreads), "row tile");
    const int t = threadIdx.x;
    int rr, p;
    bool lane_on = true;
    if constexpr (P <= 64) {
        constexpr int W = 64 / P;
        const int w = t >> 6, l = t & 63;
        const int rw = l / P;
        p = l - rw * P;
        lane_on = rw < W;
        rr = lane_on ? w * W + rw : RT + w;  // RT + w: the wavefront's dummy row
    } else {
        rr = t / P;
        p = t % P;
    }
    const int64_t row = blk * RT + rr;
    const bool valid = lane_on && row < a.nrows;
    const int64_t rowc = valid ? row : a.nrows - 1;
    const int lbase = rr * LP;
    float2* __restrict__ sp = a.spec + rowc * HP;
    const float2* __restrict__ rt = a.rtw + p;
    const int64_t r2 = rowc * L;  // real row start: float2 pairs (even rows) or floats (ODD)
    // real-space value(s) m of the row: a pair, or one value in .x (ODD)
    auto ldr = [&](const float* __restrict__ base, int m) -> float2 {
        if constexpr (ODD) return make_float2(ld_f1<kNtRealLd>(base + r2 + m), 0.f);
        else return ld_f2<kNtRealLd>(reinterpret_cast<const float2*>(base) + r2 + m);
    };
    auto str = [&](float* __restrict__ base, int m, float2 o) {
        if constexpr (ODD) st_f1<kNtRealSt>(base + r2 + m, o.x);
        else st_f2<kNtRealSt>(reinterpret_cast<float2*>(base) + r2 + m, o);
    };
    // the same for I16 (integers in [0, 65535]: exact either way)
    auto ld16 = [&](const uint16_t* __restrict__ base, int m) -> float2 {
        if constexpr (ODD) {
            return make_float2((float)base[r2 + m], 0.f);
        } else {
            const uint32_t w = reinterpret_cast<const uint32_t*>(base)[r2 + m];
            return make_float2((float)(w & 0xffffu), (float)(w >> 16));
        }
    };
    auto st16 = [&](uint16_t* __restrict__ base, int m, float2 o) {
        if constexpr (ODD) base[r2 + m] = (uint16_t)o.x;
        else reinterpret_cast<uint32_t*>(base)[r2 + m] = (uint32_t)o.x | ((uint32_t)o.y << 16);
    };

    RawRow rr_src{0, false};
    if constexpr (MODE == XM_PREP_U16 || MODE == XM_PREP_F32) {
        if (a.s0) rr_src = raw_row(a, rowc, a.pn2);
    }
    float2 opA[E], opB[E];
    auto load_operands = [&]() {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = p + e * P;
            opA[e] = make_float2(0.f, 0.f);
            opB[e] = make_float2(0.f, 0.f);
            if constexpr (MODE == XM_R2C) opA[e] = ldr(a.src, m);
            if constexpr (MODE == XM_PREP_U16) {
                if (a.s0) {
                    opA[e] = ld_raw_padded(a.raw16, a, rr_src, ODD ? (int64_t)m : 2 * (int64_t)m, !ODD);
                } else if constexpr (ODD) {
                    opA[e] = make_float2((float)a.raw16[r2 + m], 0.f);
                } else {
                    const uint32_t w = reinterpret_cast<const uint32_t*>(a.raw16)[r2 + m];
                    opA[e] = make_float2((float)(w & 0xffffu), (float)(w >> 16));
                }
            }
            if constexpr (MODE == XM_PREP_F32)
                opA[e] = a.s0 ? ld_raw_padded(a.rawf, a, rr_src, ODD ? (int64_t)m : 2 * (int64_t)m, !ODD)
                              : ldr(a.rawf, m);
            if constexpr (MODE == XM_RATIO) {
                if constexpr (xm_u16(MODE_)) opA[e] = ld16(a.I16, m);
                else opA[e] = ldr(a.I, m);
            }
            if constexpr (MODE == XM_OMW_UPDATE) {
                if constexpr (xm_u16(MODE_)) opA[e] = ld16(a.I16, m);
                else opA[e] = ldr(a.Jin ? a.Jin : a.J, m);
            }
            if constexpr (MODE == XM_BIGGS_Y) {
                opA[e] = ldr(a.J, m);
                opB[e] = ldr(a.J3, m);
            }
            if constexpr (MODE == XM_BIGGS_UPDATE) {
                opA[e] = ldr(a.Y, m);
                if (!EPI && a.acc) opB[e] = ldr(a.J4, m);
            }
        }
    };
    // real-space operands: loaded before the inverse transform (their latency
    // overlaps it) unless kXrowLateOperands, which trades that overlap for
    // fewer live registers across the transform
    if constexpr (!(IN && kXrowLateOperands)) load_operands();
    float2 v[E];
    if constexpr (IN) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = p + e * P;
            if constexpr (ODD) v[e] = m < H1 ? ld_f2<kNtRowLd>(sp + m) : cconj(ld_f2<kNtRowLd>(sp + (L - m)));  // the Hermitian half supplies the rest
            else v[e] = c2r_pre(ld_f2<kNtRowLd>(sp + m), cconj(ld_f2<kNtRowLd>(sp + (L - m))), rt[e * P]);
        }
        ct_fft<L, E, true>(v, p, lds, lbase, tw);
        if constexpr (kXrowLateOperands) load_operands();
    }

    double s_num = 0.0, s_den = 0.0;
    float lam = 0.f;
    if constexpr (MODE == XM_BIGGS_Y) lam = *a.lambda;
    EpiRow er{0, 0, false};
    if constexpr (EPI) {
        if (valid) er = epi_row(a.ep, row);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = p + e * P;
        float2 o = make_float2(0.f, 0.f);
        if constexpr (MODE == XM_C2R) {
            if (valid) str(a.dst, m, v[e]);
        } else {
            o = x_elem<MODE>(v[e], opA[e], opB[e], a.bg, a.eps, lam);
        }
        if (valid) {
            if constexpr (MODE == XM_PREP_U16 || MODE == XM_PREP_F32) {
                if (MODE == XM_PREP_U16 && a.I16) st16(a.I16, m, o);
                else str(a.I, m, o);
                if (a.J) str(a.J, m, o);
            } else if constexpr (MODE == XM_OMW_UPDATE) {
                if constexpr (EPI) {
                    if (er.keep) epi_put(a.ep, er, ODD ? (int64_t)m : 2 * (int64_t)m, !ODD, o);
                } else {
                    str(a.J, m, o);
                }
            } else if constexpr (MODE == XM_BIGGS_Y) {
                str(a.Y, m, o);
            } else if constexpr (MODE == XM_BIGGS_UPDATE) {
                const float2 y = opA[e];
                if constexpr (EPI) {
                    if (er.keep) epi_put(a.ep, er, ODD ? (int64_t)m : 2 * (int64_t)m, !ODD, o);
                } else {
                    str(a.Jn, m, o);
                }
                const float2 d = make_float2(o.x - y.x, o.y - y.y);
                if (!EPI && a.acc) {
                    const float2 j4 = opB[e];
                    s_num += (double)(d.x * j4.x);
                    s_den += (double)(j4.x * j4.x);
                    if constexpr (!ODD) {
                        s_num += (double)(d.y * j4.y);
                        s_den += (double)(j4.y * j4.y);
                    }
                }
                if (!EPI && a.write_j4) str(a.J4, m, d);
            }
        }
        if constexpr (ODD) o.y = 0.f;  // the forward transform input is real
        v[e] = o;
    }

    bool out_spec = OUT_ALWAYS;
    if constexpr (OUT_MAYBE) out_spec = !EPI && a.r2c_out != 0;
    if (out_spec && ODD) {
        ct_fft<L, E, false>(v, p, lds, lbase, tw);
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (p + e * P < H1) st_f2<kNtRowSt>(sp + (p + e * P), v[e]);
        }
    } else if (out_spec) {
        ct_fft<L, E, false>(v, p, lds, lbase, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) lds[lds_slot<true>(lbase, p + e * P, 1)] = v[e];
        row_sync<P>();
        float2 X[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int k = p + e * P;
            const float2 A = lds[lds_slot<true>(lbase, k, 1)];
            const float2 B = cconj(lds[lds_slot<true>(lbase, k == 0 ? 0 : L - k, 1)]);
            X[e] = r2c_post(A, B, rt[e * P]);
        }
        const float2 z0 = lds[lbase];
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; ++e) st_f2<kNtRowSt>(sp + (p + e * P), X[e]);
            if (p == 0) st_f2<kNtRowSt>(sp + L, make_float2(z0.x - z0.y, 0.f));
        }
    }

    if constexpr (MODE == XM_BIGGS_UPDATE && !EPI) {
        if (a.acc) {
            __syncthreads();
            double* red = reinterpret_cast<double*>(lds);
            red[2 * t] = s_num;
            red[2 * t + 1] = s_den;
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if (t < w) {
                    red[2 * t] += red[2 * (t + w)];
                    red[2 * t + 1] += red[2 * (t + w) + 1];
                }
                __syncthreads();
            }
            if (t == 0) {
                a.partials[2 * blk] = red[0];
                a.partials[2 * blk + 1] = red[1];
            }
        }
    }
}

template <int L, int E, int MODE, bool ODD = false>
__global__ void __launch_bounds__(kMaxThreads) k_xrow(XArgs a, const float2* __restrict__ tw) {
    extern __shared__ float2 lds[];
    xrow_body<L, E, MODE, ODD>(a, tw, lds, blockIdx.x);
}

// One kernel per (E, mode) so each mode gets its own register allocation.
template <int E, int MODE>
__global__ void __launch_bounds__(kMaxThreads) k_xpass(XArgs a, FftPlan pl) {
    extern __shared__ float2 lds[];
    xpass_body<E, MODE>(a, pl, lds, blockIdx.x);
}

// ------------------------------------------------------------------ batched launches
// Many volumes ("chunks") of different sizes in one launch: block b belongs to
// the chunk c with starts[c] <= b < starts[c+1]; its arguments and transform
// plan come from the chunk's descriptor (block-uniform loads).
struct XChunk {
    XArgs a;
    FftPlan pl;
};

__device__ __forceinline__ int chunk_of(const int64_t* __restrict__ starts, int n, int64_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int E, int MODE>
__global__ void __launch_bounds__(kMaxThreads) k_xpass_batch(const XChunk* __restrict__ ch,
                                                             const int64_t* __restrict__ starts, int n) {
    extern __shared__ float2 lds[];
    const int c = chunk_of(starts, n, blockIdx.x);
    const XArgs a = ch[c].a;
    const FftPlan& pl = ch[c].pl;  // read in place: a local copy indexed by stage would live in scratch
    xpass_body<E, MODE>(a, pl, lds, (int64_t)blockIdx.x - starts[c]);
}

// Batched launch of the compile-time row kernel: every chunk of the group has
// rows of length 2L (same instantiation, same arithmetic as a single call).
template <int L, int E, int MODE, bool ODD = false>
__global__ void __launch_bounds__(kMaxThreads) k_xrow_batch(const XChunk* __restrict__ ch,
                                                            const int64_t* __restrict__ starts, int n) {
    extern __shared__ float2 lds[];
    const int c = chunk_of(starts, n, blockIdx.x);
    const XArgs a = ch[c].a;
    xrow_body<L, E, MODE, ODD>(a, ch[c].pl.tw, lds, (int64_t)blockIdx.x - starts[c]);
}

// ------------------------------------------------------------------ strided axes
enum CMode : int { CM_FWD = 0, CM_INV = 1, CM_CONV = 2, CM_CONVJ = 3 };

struct ColArgs {
    float2* data;
    const float2* otf;   // same layout as data (CM_CONV / CM_CONVJ)
    int64_t acount;      // columns along grid.x (contiguous index)
    int64_t amod;        // >0: column c is valid only if c % amod < alim (pad columns)
    int64_t alim;
    int64_t bstride;     // offset between grid.y slices
    int64_t stride;      // element stride along the transform
    int T;               // columns per block
    int mode;
    float scale;         // applied with the OTF product
    int xcd;             // >0: XCD-aware tile order (blocks go round-robin to the 8 XCDs; with this each
                         // XCD walks a contiguous range of tiles, sharing pages and L2 lines)
};

struct ColChunk {
    ColArgs a;
    FftPlan pl;
    int64_t gx;  // blocks along x of the chunk's own grid
};

// MIXED: the plan may hold stages whose radix does not divide E (compiled as
// kernels of their own, so the all-register plans keep their register budget).
// OTF_NT: load the OTF nontemporally (single volumes; the batched kernels load
// it plainly: 64 chunks ran 0.5-0.7 % faster that way, interleaved runs)
template <int E, int MODE, bool MIXED = false, bool OTF_NT = kNtOtfLd>
__device__ __forceinline__ void colpass_body(const ColArgs& a, const FftPlan& pl, float2* lds, int64_t bx,
                                             int64_t by) {
    const int t = threadIdx.x;
    const int T = a.T;
    const int p = t / T;
    const int c = t - p * T;
    const int P = pl.P;
    const bool act = p < P;  // owns an LDS slot
    const int64_t col = bx * T + c;
    const bool valid = act && (col < a.acount) && (a.amod == 0 || col % a.amod < a.alim);
    // per-lane base pointer plus a block-uniform step: addresses are
    // recomputed from (ptr0, e*step) instead of held live between load and store
    const int64_t step = (int64_t)P * a.stride;
    const int64_t off0 = col + by * a.bstride + (int64_t)p * a.stride;
    float2* __restrict__ ptr0 = a.data + (valid ? off0 : 0);
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = valid ? ld_f2<kNtColLd>(ptr0 + e * step) : make_float2(0.f, 0.f);
    if constexpr (MODE == CM_INV) {
        fft_run<E, true, false, MIXED>(v, pl, p, lds, c, T, act);
    } else if constexpr (MODE == CM_FWD) {
        fft_run<E, false, false, MIXED>(v, pl, p, lds, c, T, act);
    } else {
        // E <= 20: the OTF values are prefetched before the forward transform,
        // which hides their latency behind it.  E >= 24: they are read after it,
        // as each is used -- held through the transform they cost 2E registers,
        // which spilled the E = 32 kernels (256 VGPRs) and left the others at two
        // waves per SIMD.  Measured: 2048x2048x512's z_conv (E = 32) 3.62 ->
        // 4.39 TB/s read late; 512x512x256's (E = 16) 0.174 -> 0.197 ms read late.
        constexpr bool late = E >= 24;
        const float2* __restrict__ h0 = a.otf + (valid ? off0 : 0);
        float2 h[late ? 1 : E];
        if constexpr (!late) {
#pragma unroll
            for (int e = 0; e < E; ++e) h[e] = valid ? ld_f2<OTF_NT>(h0 + e * step) : make_float2(0.f, 0.f);
        }
        fft_run<E, false, false, MIXED>(v, pl, p, lds, c, T, act);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            float2 he;
            if constexpr (late) he = valid ? ld_f2<OTF_NT>(h0 + e * step) : make_float2(0.f, 0.f);
            else he = h[e];
            const float2 hs = make_float2(he.x * a.scale, he.y * a.scale);
            v[e] = (MODE == CM_CONVJ) ? cmulc(v[e], hs) : cmul(v[e], hs);
        }
        fft_run<E, true, false, MIXED>(v, pl, p, lds, c, T, act);
    }
    if (valid) {
        // nontemporal stores only for whole 128-byte row segments: with
        // 8-column tiles (64 bytes) they left the two halves of a line unmerged
        // (WRITE_SIZE of the 2048-point dim-2 passes +16 %)
        if (kNtColSt && T >= 16) {
#pragma unroll
            for (int e = 0; e < E; ++e) st_f2<true>(ptr0 + e * step, v[e]);
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) st_f2<false>(ptr0 + e * step, v[e]);
        }
    }
}

// NT: the block size the kernel is compiled for (256; 512 / 1024 for long
// columns, whose 8-16-column tiles need more threads and up to 160 KB of LDS).
template <int E, int MODE, int NT = kMaxThreads, bool MIXED = false>
__global__ void __launch_bounds__(NT) k_colpass(ColArgs a, FftPlan pl) {
    extern __shared__ float2 lds[];
    int64_t bx = blockIdx.x, by = blockIdx.y;
    if (a.xcd) {  // block b runs on XCD b % 8: give XCD k the k-th eighth of the tiles, in order
        const int64_t gx = gridDim.x, total = gx * gridDim.y;
        const int64_t b = bx + by * gx;
        const int64_t t = (b & 7) * (total >> 3) + (b >> 3);
        bx = t % gx;
        by = t / gx;
    }
    colpass_body<E, MODE, MIXED>(a, pl, lds, bx, by);
}

template <int E, int MODE, int NT = kMaxThreads, bool MIXED = false>
__global__ void __launch_bounds__(NT) k_colpass_batch(const ColChunk* __restrict__ ch,
                                                      const int64_t* __restrict__ starts, int n) {
    extern __shared__ float2 lds[];
    const int c = chunk_of(starts, n, blockIdx.x);
    const ColArgs a = ch[c].a;
    const FftPlan& pl = ch[c].pl;  // read in place: a local copy indexed by stage would live in scratch
    const int64_t lb = (int64_t)blockIdx.x - starts[c];
    colpass_body<E, MODE, MIXED, false>(a, pl, lds, lb % ch[c].gx, lb / ch[c].gx);
}

}  // namespace pk
