// pk_batch.hip -- many volumes ("chunks") of different sizes deconvolved
// together, the workload RLdecon_large_in_memory.m:190-260 /
// RLdecon_for_zarr_block.m:101-171 produce with their per-chunk loops.
//
// Every pass of the iteration runs as ONE launch per values-per-thread group
// over all chunks: a device descriptor table holds each chunk's arguments and
// transform plan, and each block finds its chunk by a binary search over the
// chunks' block offsets.  Each chunk keeps the exact tile shape and plan of
// the single-volume path, so its arithmetic -- including the block partial
// sums of the Biggs step length -- is bitwise the same alone or in a batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "pk_engine.hpp"
#include "pk_launch.hpp"
#include "pk_misc.hpp"

namespace pk {

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

template <class T>
struct Group {
    int E = 0;
    std::vector<T> chunks;
    std::vector<int64_t> starts;  // block offsets, size chunks+1
    unsigned block = 1;
    size_t lds = 16;
    size_t dev_off = 0;   // byte offset of chunks in the table buffer
    size_t start_off = 0; // byte offset of starts
};

// One pass over the batch: groups keyed by E.
template <class T>
using Pass = std::map<int, Group<T>>;

struct BItem {
    const VolPlan* vp = nullptr;
    int64_t d[3] = {0, 0, 0};  // caller's size
    const float2* Hf = nullptr;
    const float2* Hb = nullptr;
    float *I = nullptr, *J = nullptr, *J3 = nullptr, *Y = nullptr, *J4 = nullptr, *lam = nullptr;
    float2* A = nullptr;
    double* part = nullptr;
    const uint16_t* raw16 = nullptr;
    const float* rawf = nullptr;
    void* dout = nullptr;  // device output (caller's or staging)
    int64_t od[3] = {0, 0, 0};
    int64_t b0[3] = {0, 0, 0};
};

}  // namespace

__global__ void k_pad_copy(const void* __restrict__ src, void* __restrict__ dst, int esz, int64_t s1, int64_t s2,
                           int64_t s3, int64_t d1, int64_t d2, int64_t d3) {
    const int64_t n = d1 * d2 * d3;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t x = i % d1, y = (i / d1) % d2, z = i / (d1 * d2);
        const bool in = x < s1 && y < s2 && z < s3;
        const int64_t si = x + s1 * (y + s2 * z);
        if (esz == 2)
            static_cast<uint16_t*>(dst)[i] = in ? static_cast<const uint16_t*>(src)[si] : (uint16_t)0;
        else
            static_cast<float*>(dst)[i] = in ? static_cast<const float*>(src)[si] : 0.f;
    }
}

// Per-chunk Biggs step length: block c reduces chunk c's partials in the same
// fixed order as the single-volume k_lambda.
struct LamChunk {
    const double* partials;
    int nblk;
    float* lambda;
};
__global__ void k_lambda_batch(const LamChunk* __restrict__ ch, float eps) {
    __shared__ double sn[256], sd[256];
    const LamChunk c = ch[blockIdx.x];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int i = t; i < c.nblk; i += 256) {
        a += c.partials[2 * i];
        b += c.partials[2 * i + 1];
    }
    sn[t] = a;
    sd[t] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            sn[t] += sn[t + w];
            sd[t] += sd[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const float num = (float)sn[0];
        const float den = (float)sd[0];
        float l = num / (den + eps);
        l = (l > 1.f || l != l) ? 1.f : l;
        l = l > 0.f ? l : 0.f;
        *c.lambda = l;
    }
}

static int64_t smooth_up(int64_t n) {
    while (!smooth235(n)) ++n;
    return n;
}

void pk_ctx_impl::decon_batch(int method, const pk_batch_item* items, int n, const float* psf_f, const int64_t fd[3],
                              const float* psf_b, const int64_t bd[3], int numit, const pk_opts* o) {
    static const float kEpsF = 2.220446049250313080847e-16f;
    if (!items || n < 1 || !o || !psf_f || numit < 0) fail(PK_ERR_INVALID_ARG, "empty batch or null argument");
    if (method == 1 && !psf_b) fail(PK_ERR_INVALID_ARG, "PSF_b missing");
    if (o->has_bbox) fail(PK_ERR_INVALID_ARG, "bbox is per chunk: crop the outputs instead");
    std::vector<BItem> it((size_t)n);
    // ---- plans, PSF crops (RLdecon.m:264-270), OTFs
    std::map<std::vector<int64_t>, int> distinct;
    for (int i = 0; i < n; ++i) {
        const pk_vol& v = items[i].in;
        if (v.dtype != PK_UINT16 && v.dtype != PK_SINGLE) fail(PK_ERR_UNSUPPORTED_TYPE, "input class must be uint16 or single");
        if (!v.data || !items[i].out.data) fail(PK_ERR_INVALID_ARG, "null data pointer");
        for (int d = 0; d < 3; ++d) {
            if (v.dims[d] < 1) fail(PK_ERR_INVALID_ARG, "empty chunk");
            it[(size_t)i].d[d] = v.dims[d];
            it[(size_t)i].od[d] = v.dims[d];
        }
        const pk_vol& ov = items[i].out;
        if (ov.dims[0] != v.dims[0] || ov.dims[1] != v.dims[1] || ov.dims[2] != v.dims[2])
            fail(PK_ERR_INVALID_ARG, "output dims must equal input dims");
        if (ov.dtype != (o->save16bit ? PK_UINT16 : PK_SINGLE)) fail(PK_ERR_INVALID_ARG, "output class mismatch");
        distinct[{smooth_up(v.dims[0]), smooth_up(v.dims[1]), smooth_up(v.dims[2])}]++;
    }
    otf_max_entries = std::max<size_t>(otf_max_entries, 2 * distinct.size() + 4);

    auto crop = [](const float* p, const int64_t pd[3], const int64_t im[3], std::vector<float>& out, int64_t cd[3]) {
        int64_t s[3];
        for (int d = 0; d < 3; ++d) {
            s[d] = pd[d] > im[d] ? (pd[d] - im[d]) / 2 : 0;
            cd[d] = std::min(pd[d], im[d]);
        }
        out.resize((size_t)(cd[0] * cd[1] * cd[2]));
        for (int64_t k = 0; k < cd[2]; ++k)
            for (int64_t j = 0; j < cd[1]; ++j)
                for (int64_t i = 0; i < cd[0]; ++i)
                    out[(size_t)(i + cd[0] * (j + cd[1] * k))] = p[(i + s[0]) + pd[0] * ((j + s[1]) + pd[1] * (k + s[2]))];
    };

    size_t arena = 0;
    std::vector<size_t> offs((size_t)n);
    for (int i = 0; i < n; ++i) {
        BItem& b = it[(size_t)i];
        const int64_t pd3[3] = {smooth_up(b.d[0]), smooth_up(b.d[1]), smooth_up(b.d[2])};
        b.vp = &plan(pd3[0], pd3[1], pd3[2]);
        // PSF larger than the chunk: centred crop as RLdecon does per chunk
        std::vector<float> cf, cb;
        int64_t cfd[3], cbd[3];
        crop(psf_f, fd, b.d, cf, cfd);
        b.Hf = otf_for(*b.vp, cf.data(), cfd);
        if (method == 1) {
            crop(psf_b, bd, b.d, cb, cbd);
            b.Hb = otf_for(*b.vp, cb.data(), cbd);
        }
        const VolPlan& vp = *b.vp;
        const bool padded = pd3[0] != b.d[0] || pd3[1] != b.d[1] || pd3[2] != b.d[2];
        const size_t esz = items[i].in.dtype == PK_UINT16 ? 2 : 4;
        const int64_t nxb = (vp.n2 * vp.n3 + vp.x.tile - 1) / vp.x.tile;
        size_t need = 2 * align_up(sizeof(float) * vp.N) + align_up(sizeof(float2) * vp.S);
        if (method == 0) need += 3 * align_up(sizeof(float) * vp.N) + align_up(sizeof(double) * 2 * nxb) + kAlign;
        if (padded || items[i].in.mem == PK_MEM_HOST) need += align_up(esz * vp.N);
        if (items[i].out.mem == PK_MEM_HOST) need += align_up((o->save16bit ? 2 : 4) * (size_t)(b.d[0] * b.d[1] * b.d[2]));
        offs[(size_t)i] = arena;
        arena += need;
    }
    ensure(ws.batch, arena);
    char* base = ws.batch.as<char>();
    for (int i = 0; i < n; ++i) {
        BItem& b = it[(size_t)i];
        const VolPlan& vp = *b.vp;
        char* p = base + offs[(size_t)i];
        auto take = [&](size_t bytes) {
            char* r = p;
            p += align_up(bytes);
            return r;
        };
        b.I = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
        b.J = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
        b.A = reinterpret_cast<float2*>(take(sizeof(float2) * vp.S));
        if (method == 0) {
            const int64_t nxb = (vp.n2 * vp.n3 + vp.x.tile - 1) / vp.x.tile;
            b.J3 = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
            b.Y = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
            b.J4 = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
            b.part = reinterpret_cast<double*>(take(sizeof(double) * 2 * nxb));
            b.lam = reinterpret_cast<float*>(take(kAlign));
        }
        const pk_vol& v = items[i].in;
        const bool u16 = v.dtype == PK_UINT16;
        const size_t esz = u16 ? 2 : 4;
        const bool padded = vp.n1 != b.d[0] || vp.n2 != b.d[1] || vp.n3 != b.d[2];
        const void* raw = v.data;
        if (padded || v.mem == PK_MEM_HOST) {
            void* dst = take(esz * vp.N);
            const void* src = v.data;
            if (v.mem == PK_MEM_HOST) {
                // stage the caller's chunk (unpadded) right behind the padded buffer's start
                PK_HIP(hipMemcpyAsync(padded ? reinterpret_cast<char*>(b.J) : dst, v.data,
                                      esz * (size_t)(b.d[0] * b.d[1] * b.d[2]), hipMemcpyHostToDevice, stream));
                src = padded ? static_cast<const void*>(b.J) : dst;
            }
            if (padded) {
                hipLaunchKernelGGL(k_pad_copy, dim3((unsigned)std::min<int64_t>((vp.N + 255) / 256, 65536)), dim3(256), 0,
                                   stream, src, dst, (int)esz, b.d[0], b.d[1], b.d[2], vp.n1, vp.n2, vp.n3);
                PK_HIP(hipGetLastError());
            }
            raw = dst;
        }
        b.raw16 = u16 ? static_cast<const uint16_t*>(raw) : nullptr;
        b.rawf = u16 ? nullptr : static_cast<const float*>(raw);
        b.dout = items[i].out.mem == PK_MEM_HOST ? static_cast<void*>(take((o->save16bit ? 2 : 4) *
                                                                           (size_t)(b.d[0] * b.d[1] * b.d[2])))
                                                 : items[i].out.data;
    }

    // ---- descriptor tables, one per distinct pass configuration
    std::vector<char> table;
    auto upload_pos = [&](size_t bytes) {
        const size_t off = align_up(table.size());
        table.resize(off + bytes);
        return off;
    };
    auto base_x = [&](const BItem& b, int mode) {
        XArgs a{};
        const VolPlan& vp = *b.vp;
        a.n1 = (int)vp.n1;
        a.h1 = (int)vp.h1;
        a.hp = (int)vp.hp;
        a.packed = vp.packed;
        a.RT = vp.x.tile;
        a.lpad = vp.x.lpad;
        a.mode = mode;
        a.nrows = vp.n2 * vp.n3;
        a.eps = kEpsF;
        a.rtw = vp.rtw;
        a.spec = b.A;
        return a;
    };
    auto build_x = [&](auto&& make) {  // make(item) -> XArgs
        Pass<XChunk> ps;
        for (auto& b : it) {
            const VolPlan& vp = *b.vp;
            Group<XChunk>& g = ps[vp.x.fp.E];
            g.E = vp.x.fp.E;
            if (g.starts.empty()) g.starts.push_back(0);
            g.chunks.push_back(XChunk{make(b), vp.x.fp});
            g.starts.push_back(g.starts.back() + (vp.n2 * vp.n3 + vp.x.tile - 1) / vp.x.tile);
            g.block = std::max(g.block, vp.x.block.x);
            g.lds = std::max(g.lds, vp.x.lds);
        }
        for (auto& kv : ps) {
            auto& g = kv.second;
            g.dev_off = upload_pos(sizeof(XChunk) * g.chunks.size());
            std::memcpy(table.data() + g.dev_off, g.chunks.data(), sizeof(XChunk) * g.chunks.size());
            g.start_off = upload_pos(sizeof(int64_t) * g.starts.size());
            std::memcpy(table.data() + g.start_off, g.starts.data(), sizeof(int64_t) * g.starts.size());
        }
        return ps;
    };
    auto build_col = [&](int axis, int mode, bool use_hb) {
        Pass<ColChunk> ps;
        for (auto& b : it) {
            const VolPlan& vp = *b.vp;
            const AxisLaunch& ax = axis == 2 ? vp.y : vp.z;
            if (ax.fp.L == 1 && (mode == CM_FWD || mode == CM_INV)) continue;
            ColArgs a{};
            a.data = b.A;
            a.otf = use_hb ? b.Hb : b.Hf;
            a.T = ax.tile;
            a.mode = mode;
            a.scale = (float)(1.0 / (double)vp.N);
            int64_t gx, gy;
            if (axis == 2) {
                a.acount = vp.h1;
                a.bstride = vp.hp * vp.n2;
                a.stride = vp.hp;
                gx = (vp.h1 + a.T - 1) / a.T;
                gy = vp.n3;
            } else {
                a.acount = vp.hp * vp.n2;
                a.amod = vp.hp;
                a.alim = vp.h1;
                a.stride = vp.hp * vp.n2;
                gx = (a.acount + a.T - 1) / a.T;
                gy = 1;
            }
            Group<ColChunk>& g = ps[ax.fp.E];
            g.E = ax.fp.E;
            if (g.starts.empty()) g.starts.push_back(0);
            g.chunks.push_back(ColChunk{a, ax.fp, gx});
            g.starts.push_back(g.starts.back() + gx * gy);
            g.block = std::max(g.block, ax.block.x);
            g.lds = std::max(g.lds, ax.lds);
        }
        for (auto& kv : ps) {
            auto& g = kv.second;
            g.dev_off = upload_pos(sizeof(ColChunk) * g.chunks.size());
            std::memcpy(table.data() + g.dev_off, g.chunks.data(), sizeof(ColChunk) * g.chunks.size());
            g.start_off = upload_pos(sizeof(int64_t) * g.starts.size());
            std::memcpy(table.data() + g.start_off, g.starts.data(), sizeof(int64_t) * g.starts.size());
        }
        return ps;
    };

    const bool raw_u16 = items[0].in.dtype == PK_UINT16;
    for (int i = 1; i < n; ++i)
        if ((items[i].in.dtype == PK_UINT16) != raw_u16) fail(PK_ERR_INVALID_ARG, "all chunks must share one input class");
    const float bg = raw_u16 ? (float)std::min(65535.0, std::max(0.0, std::round(o->background)))
                             : (float)o->background;
    auto prep = build_x([&](const BItem& b) {
        XArgs a = base_x(b, raw_u16 ? XM_PREP_U16 : XM_PREP_F32);
        a.raw16 = b.raw16;
        a.rawf = b.rawf;
        a.bg = bg;
        a.I = b.I;
        a.J = b.J;
        return a;
    });
    auto yf = build_col(2, CM_FWD, false);
    auto yi = build_col(2, CM_INV, false);
    auto zf = build_col(3, CM_CONV, false);
    auto zb = build_col(3, method == 1 ? CM_CONV : CM_CONVJ, method == 1);
    auto ratio = build_x([&](const BItem& b) {
        XArgs a = base_x(b, XM_RATIO);
        a.I = b.I;
        return a;
    });
    Pass<XChunk> upd_more, upd_last;
    if (method == 1) {
        upd_more = build_x([&](const BItem& b) {
            XArgs a = base_x(b, XM_OMW_UPDATE);
            a.J = b.J;
            a.r2c_out = 1;
            return a;
        });
        upd_last = build_x([&](const BItem& b) {
            XArgs a = base_x(b, XM_OMW_UPDATE);
            a.J = b.J;
            a.r2c_out = 0;
            return a;
        });
    }
    std::vector<LamChunk> lam;
    for (auto& b : it) {
        if (method == 0)
            lam.push_back(LamChunk{b.part, (int)((b.vp->n2 * b.vp->n3 + b.vp->x.tile - 1) / b.vp->x.tile), b.lam});
    }
    size_t lam_off = 0;
    if (method == 0) {
        lam_off = upload_pos(sizeof(LamChunk) * lam.size());
        std::memcpy(table.data() + lam_off, lam.data(), sizeof(LamChunk) * lam.size());
    }
    // Biggs passes depend on which buffer holds x_{k-1} (the estimates
    // ping-pong between J and J3): build both parities; the update comes in
    // three flavours (acc, write_j4) = (0,1), (1,1), (0,0)
    static const int kUpd[3][2] = {{0, 1}, {1, 1}, {0, 0}};
    Pass<XChunk> biggs_y[2], biggs_up[2][3];
    if (method == 0) {
        for (int par = 0; par < 2; ++par) {
            biggs_y[par] = build_x([&](const BItem& b) {
                XArgs a = base_x(b, XM_BIGGS_Y);
                a.J = par ? b.J3 : b.J;
                a.J3 = par ? b.J : b.J3;
                a.Y = b.Y;
                a.lambda = b.lam;
                return a;
            });
            for (int u = 0; u < 3; ++u)
                biggs_up[par][u] = build_x([&](const BItem& b) {
                    XArgs a = base_x(b, XM_BIGGS_UPDATE);
                    a.Y = b.Y;
                    a.Jn = par ? b.J : b.J3;
                    a.J4 = b.J4;
                    a.acc = kUpd[u][0];
                    a.write_j4 = kUpd[u][1];
                    a.partials = b.part;
                    return a;
                });
        }
    }
    ensure(ws.batch_tab, table.size());
    PK_HIP(hipMemcpyAsync(ws.batch_tab.p, table.data(), table.size(), hipMemcpyHostToDevice, stream));
    char* tab = ws.batch_tab.as<char>();

    auto run_xp = [&](const Pass<XChunk>& ps, int mode, const char* name) {
        for (auto& kv : ps) {
            const auto& g = kv.second;
            hipEvent_t e0 = prof_mark();
            PK_HIP(launch_xpass_batch(g.E, dim3((unsigned)g.starts.back()), dim3(g.block), g.lds, stream, mode,
                                      reinterpret_cast<const XChunk*>(tab + g.dev_off),
                                      reinterpret_cast<const int64_t*>(tab + g.start_off), (int)g.chunks.size()));
            prof_done(name, 0, e0);
        }
    };
    auto run_cp = [&](const Pass<ColChunk>& ps, int mode, const char* name) {
        for (auto& kv : ps) {
            const auto& g = kv.second;
            hipEvent_t e0 = prof_mark();
            PK_HIP(launch_colpass_batch(g.E, dim3((unsigned)g.starts.back()), dim3(g.block), g.lds, stream, mode,
                                        reinterpret_cast<const ColChunk*>(tab + g.dev_off),
                                        reinterpret_cast<const int64_t*>(tab + g.start_off), (int)g.chunks.size()));
            prof_done(name, 0, e0);
        }
    };

    // ---- the iteration schedule (same order of passes as the single-volume path)
    run_xp(prep, raw_u16 ? XM_PREP_U16 : XM_PREP_F32, "b_x_prep");
    int par = 0;  // Biggs: 0 -> x_{k-1} in J, x_{k-2} in J3
    if (method == 0) {
        for (auto& b : it) {
            hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)std::min<int64_t>((b.vp->N + 255) / 256, 65536)), dim3(256), 0,
                               stream, b.J3, b.vp->N, 0.f);
            hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, stream, b.lam, (int64_t)1, 0.f);
        }
        PK_HIP(hipGetLastError());
    }
    for (int k = 1; k <= numit; ++k) {
        if (method == 0) {
            if (k > 2) {
                hipLaunchKernelGGL(k_lambda_batch, dim3((unsigned)n), dim3(256), 0, stream,
                                   reinterpret_cast<const LamChunk*>(tab + lam_off), kEpsF);
                PK_HIP(hipGetLastError());
            }
            run_xp(biggs_y[par], XM_BIGGS_Y, "b_x_biggs_y");
        }
        run_cp(yf, CM_FWD, "b_y_fwd");
        run_cp(zf, CM_CONV, "b_z_conv");
        run_cp(yi, CM_INV, "b_y_inv");
        run_xp(ratio, XM_RATIO, "b_x_ratio");
        run_cp(yf, CM_FWD, "b_y_fwd");
        run_cp(zb, method == 1 ? CM_CONV : CM_CONVJ, "b_z_conv");
        run_cp(yi, CM_INV, "b_y_inv");
        if (method == 1) {
            run_xp(k < numit ? upd_more : upd_last, XM_OMW_UPDATE, "b_x_update");
        } else {
            const int u = (k >= 2 && k < numit) ? 1 : (k < numit ? 0 : 2);
            run_xp(biggs_up[par][u], XM_BIGGS_UPDATE, "b_x_biggs_update");
            par ^= 1;
        }
    }

    // ---- epilogue and output per chunk (decon_lucy_function.m:177-203)
    for (int i = 0; i < n; ++i) {
        BItem& b = it[(size_t)i];
        const VolPlan& vp = *b.vp;
        EpiArgs e{};
        e.J = (method == 0 && par == 1) ? b.J3 : b.J;
        e.I = b.I;
        e.raw16 = b.raw16;
        e.rawf = b.rawf;
        e.out16 = o->save16bit ? static_cast<uint16_t*>(b.dout) : nullptr;
        e.outf = o->save16bit ? nullptr : static_cast<float*>(b.dout);
        e.n1 = vp.n1;
        e.n2 = vp.n2;
        e.o1 = b.od[0];
        e.o2 = b.od[1];
        e.o3 = b.od[2];
        e.damp = o->damp_factor > 1 ? (float)o->damp_factor : 0.f;
        e.scale = (float)o->scale_factor;
        e.offset = (float)o->decon_offset;
        e.apply_scale = (o->scale_factor != 1 || o->decon_offset != 0 || o->edge_erosion > 0);
        e.use_mask = (o->decon_offset != 0 || o->edge_erosion > 0);
        if (o->edge_erosion > 0) fail(PK_ERR_INVALID_ARG, "EdgeErosion is not supported in batches");
        const int64_t on = b.od[0] * b.od[1] * b.od[2];
        hipLaunchKernelGGL(k_epilogue, dim3((unsigned)std::min<int64_t>((on + 255) / 256, 65536)), dim3(256), 0, stream, e);
        PK_HIP(hipGetLastError());
        if (items[i].out.mem == PK_MEM_HOST)
            PK_HIP(hipMemcpyAsync(items[i].out.data, b.dout, (o->save16bit ? 2 : 4) * (size_t)on, hipMemcpyDeviceToHost,
                                  stream));
    }
    PK_HIP(hipStreamSynchronize(stream));
    prof_flush();
}

}  // namespace pk
