// pk_plain.hip -- the MATLAB-structured comparison path (PK_FLAG_UNFUSED):
// the iteration of decon_lucy_function.m:121-142 and
// decon_lucy_omw_function.m:105-112 with one kernel per gpuArray statement --
// fftn / ifftn as separate 3-pass transforms, the OTF product, max, ./ and .*
// each a full-array pass.  Same transforms, plans and arithmetic as the fused
// engine; only the fusion differs.  Used as the "plain GPU path" baseline.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pk_engine.hpp"
#include "pk_misc.hpp"

namespace pk {

enum EwOp : int {
    EW_PREP_U16 = 0,  // c = max(u16 a - bg, 0)   (I - uint16(Background), single(I))
    EW_PREP_F32,      // c = max(a - bg, 0)
    EW_MAX_EPS,       // c = max(a, eps)
    EW_DIV,           // c = a ./ b
    EW_MUL,           // c = a .* b
    EW_MAX0,          // c = max(a, 0)
    EW_SUB,           // c = a - b
    EW_EXTRAP,        // c = a + lambda * (a - b)
    EW_COPY,          // c = a
};

__global__ void k_ew(int op, const void* __restrict__ av, const float* __restrict__ b, float* __restrict__ c, int64_t n,
                     float s, const float* __restrict__ lam) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const float L = lam ? *lam : 0.f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float* a = static_cast<const float*>(av);
        float r = 0.f;
        switch (op) {
            case EW_PREP_U16: {
                const float v = (float)static_cast<const uint16_t*>(av)[i] - s;
                r = v > 0.f ? v : 0.f;
                break;
            }
            case EW_PREP_F32: {
                const float v = a[i] - s;
                r = v > 0.f ? v : 0.f;
                break;
            }
            case EW_MAX_EPS: r = a[i] > s ? a[i] : s; break;
            case EW_DIV: r = a[i] / b[i]; break;
            case EW_MUL: r = a[i] * b[i]; break;
            case EW_MAX0: r = a[i] > 0.f ? a[i] : 0.f; break;
            case EW_SUB: r = a[i] - b[i]; break;
            case EW_EXTRAP: r = a[i] + L * (a[i] - b[i]); break;
            case EW_COPY: r = a[i]; break;
        }
        c[i] = r;
    }
}

// W .* H (or conj(H)) scaled by s over a pitched half spectrum.
__global__ void k_spec_mul(float2* __restrict__ w, const float2* __restrict__ h, int64_t n, float s, int conj) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float2 a = w[i];
        float2 b = h[i];
        b.x *= s;
        b.y *= s;
        if (conj) b.y = -b.y;
        w[i] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
}

// The two Biggs dot products as MATLAB forms them (sum((J_2(:) - Y(:)) .* J_4)
// and sum(J_4 .* J_4)), per-block FP64 partials for k_lambda.
__global__ void k_dot2(const float* __restrict__ j2, const float* __restrict__ y, const float* __restrict__ j4, int64_t n,
                       double* __restrict__ part) {
    __shared__ double sa[256], sb[256];
    double a = 0.0, b = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float d = j2[i] - y[i];
        a += (double)(d * j4[i]);
        b += (double)(j4[i] * j4[i]);
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sa[0];
        part[2 * blockIdx.x + 1] = sb[0];
    }
}

void pk_ctx_impl::decon_unfused(int method, const VolPlan& vp, const void* raw, bool u16, float bg, const float2* Hf,
                                const float2* Hb, int numit, float* I, float* J) {
    static const float kE = 2.220446049250313080847e-16f;
    const int64_t N = vp.N;
    const unsigned g = (unsigned)std::min<int64_t>((N + 255) / 256, 8192);
    const float inv = (float)(1.0 / (double)N);
    ensure(ws.A, sizeof(float2) * vp.S);
    ensure(ws.tmp, sizeof(float) * N);
    float2* W = ws.A.as<float2>();
    float* T = ws.tmp.as<float>();
    auto ew = [&](int op, const void* a, const float* b, float* c, float s = 0.f, const float* lam = nullptr) {
        hipEvent_t e0 = prof_mark();
        hipLaunchKernelGGL(k_ew, dim3(g), dim3(256), 0, stream, op, a, b, c, N, s, lam);
        PK_HIP(hipGetLastError());
        prof_done("u_elementwise", (double)N * (op == EW_DIV || op == EW_MUL || op == EW_SUB || op == EW_EXTRAP ? 12 : 8), e0);
    };
    auto fftn = [&](const float* x) { r2c3d(vp, x, W); };
    auto ifftn_real = [&](float* y) { c2r3d(vp, W, y); };
    auto otf_mul = [&](const float2* H, bool conj) {
        hipEvent_t e0 = prof_mark();
        hipLaunchKernelGGL(k_spec_mul, dim3((unsigned)std::min<int64_t>((vp.S + 255) / 256, 8192)), dim3(256), 0, stream, W,
                           H, vp.S, inv, conj ? 1 : 0);
        PK_HIP(hipGetLastError());
        prof_done("u_otf_mul", 24.0 * vp.Sa, e0);
    };
    // I = I - uint16(Background); I = single(I)  |  I = max(I - Background, 0)
    ew(u16 ? EW_PREP_U16 : EW_PREP_F32, raw, nullptr, I, bg);
    ew(EW_COPY, I, nullptr, J);
    if (method == 1) {
        for (int k = 1; k <= numit; ++k) {
            fftn(J);                          // fftn(J)
            otf_mul(Hf, false);               // .* OTF_f
            ifftn_real(T);                    // CX = real(ifftn(...))
            ew(EW_MAX_EPS, T, nullptr, T, kE); // CX = max(CX, eps)
            ew(EW_DIV, I, T, T);              // I ./ CX
            fftn(T);
            otf_mul(Hb, false);               // .* OTF_b
            ifftn_real(T);                    // real(ifftn(...))
            ew(EW_MUL, T, J, J);              // ... .* J
            ew(EW_MAX0, J, nullptr, J);       // J = max(J, 0)
        }
        return;
    }
    // Biggs RL (decon_lucy_function.m:120-142)
    float* J2 = J;
    float* J3 = ws.J3.as<float>();
    float* Y = ws.Y.as<float>();
    float* J4 = ws.J4.as<float>();
    float* lam = ws.lambda.as<float>();
    double* part = ws.partials.as<double>();
    const int nblk = 1024;
    hipLaunchKernelGGL(k_fill_f32, dim3(g), dim3(256), 0, stream, J3, N, 0.f);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, stream, lam, (int64_t)1, 0.f);
    PK_HIP(hipGetLastError());
    for (int k = 1; k <= numit; ++k) {
        if (k > 2) {
            hipLaunchKernelGGL(k_dot2, dim3(nblk), dim3(256), 0, stream, J2, Y, J4, N, part);
            hipLaunchKernelGGL(k_lambda, dim3(1), dim3(256), 0, stream, part, nblk, kE, lam);
            PK_HIP(hipGetLastError());
            ew(EW_SUB, J2, Y, J4);            // J_4 = J_2(:) - Y(:)
        } else if (k == 2) {
            ew(EW_SUB, J2, Y, J4);
        }
        ew(EW_EXTRAP, J2, J3, Y, 0.f, lam);   // J_2 + lambda*(J_2 - J_3)
        ew(EW_MAX0, Y, nullptr, Y);           // Y = max(..., 0)
        fftn(Y);
        otf_mul(Hf, false);                   // H .* fftn(Y)
        ifftn_real(T);
        ew(EW_MAX_EPS, T, nullptr, T, kE);    // ReBlurred = max(real(ifftn(...)), eps)
        ew(EW_DIV, I, T, T);                  // ReBlurred = I ./ ReBlurred
        fftn(T);
        otf_mul(Hf, true);                    // conj(H) .* fftn(ReBlurred)
        ifftn_real(T);
        std::swap(J2, J3);                    // J_3 = J_2
        ew(EW_MUL, Y, T, J2);                 // Y .* real(ifftn(...))
        ew(EW_MAX0, J2, nullptr, J2);         // J_2 = max(..., 0)
    }
    if (J2 != J) ew(EW_COPY, J2, nullptr, J);
}

}  // namespace pk
