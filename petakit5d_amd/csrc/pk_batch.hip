// pk_batch.hip -- many volumes ("chunks") in flight at once: the batched
// decon of RLdecon_large_in_memory.m:190-260 / RLdecon_for_zarr_block.m:101-171
// chunk loops, and the request loop that serves decon jobs.
//
// The step engine advances every active chunk by ONE iteration.  Each pass of
// the iteration is one launch per (values-per-thread, mode) group over all
// chunks: a device descriptor table holds each chunk's arguments and
// transform plan, and each block finds its chunk by a binary search over the
// chunks' block offsets.  Chunks may sit at different iterations and mix RL
// and OMW.  Each chunk keeps the exact tile shape and plan of the
// single-volume path, so its arithmetic -- including the block partial sums
// of the Biggs step length -- is bitwise the same alone, in a batch, or in
// a served mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <array>
#include <tuple>
#include <vector>

#include "pk_engine.hpp"
#include "pk_launch.hpp"
#include "pk_misc.hpp"

namespace pk {

namespace {
constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
const float kEpsF = 2.220446049250313080847e-16f;

template <class T>
struct Group {
    int E = 0;
    int mode = 0;
    int nt = 256;  // block size the kernels are compiled for
    int L = 0;     // x passes: row length of a compile-time row kernel shared by the group (0: runtime plan);
                   // strided passes: 1 for plans with LDS-fed stages
    std::vector<T> chunks;
    std::vector<int64_t> starts{0};  // block offsets, size chunks+1
    unsigned block = 1;
    size_t lds = 16;
    size_t dev_off = 0;
    size_t start_off = 0;
};
// Chunks share a launch only when their blocks have the same shape: values per
// thread, mode, compiled block size, row length of a compile-time row kernel,
// threads per block and LDS bytes (no idle threads, no over-allocated LDS).
template <class T>
using Pass = std::map<std::array<int64_t, 6>, Group<T>>;

int64_t smooth_up(int64_t n) {
    while (!smooth235(n)) ++n;
    return n;
}
}  // namespace

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

// ------------------------------------------------------------------ chunk set-up
// Centred crop of a PSF larger than the chunk (RLdecon.m:264-270, 326-332).
static void crop_psf(const float* p, const int64_t pd[3], const int64_t im[3], std::vector<float>& out, int64_t cd[3]) {
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
}

void pk_ctx_impl::chunk_validate(const pk_vol& in, const pk_vol& out, const pk_opts& o) {
    if (in.dtype != PK_UINT16 && in.dtype != PK_SINGLE) fail(PK_ERR_UNSUPPORTED_TYPE, "input class must be uint16 or single");
    if (!in.data || !out.data) fail(PK_ERR_INVALID_ARG, "null data pointer");
    for (int d = 0; d < 3; ++d) {
        if (in.dims[d] < 1) fail(PK_ERR_INVALID_ARG, "empty volume");
        if (out.dims[d] != in.dims[d]) fail(PK_ERR_INVALID_ARG, "output dims must equal input dims");
    }
    if (out.dtype != (o.save16bit ? PK_UINT16 : PK_SINGLE)) fail(PK_ERR_INVALID_ARG, "output class mismatch");
    if (o.has_bbox) fail(PK_ERR_INVALID_ARG, "bbox is per volume: crop the output instead");
    if (o.edge_erosion > 0) fail(PK_ERR_INVALID_ARG, "EdgeErosion is not supported for batched volumes");
}

// Plans and OTFs of a chunk; returns the device bytes its state needs.
size_t pk_ctx_impl::chunk_prepare(Chunk& c, const pk_vol& in, const pk_vol& out, const float* psf_f,
                                  const int64_t fd[3], uint64_t hf, const float* psf_b, const int64_t bd[3],
                                  uint64_t hb) {
    for (int d = 0; d < 3; ++d) c.d[d] = in.dims[d];
    {
        const int64_t a = smooth_up(c.d[0]), b = smooth_up(c.d[1]), z = smooth_up(c.d[2]);
        if ((c.o.flags & PK_FLAG_AUTOTUNE) && !plan(a, b, z).tuned) autotune(a, b, z);
        c.vp = &plan(a, b, z);
    }
    // The cropped PSF is a function of the whole PSF and the crop window, so
    // its cache key is the whole PSF's key with the window folded in (the
    // whole PSF's own key when nothing is cropped); the crop is only
    // materialised to build a missing OTF.
    auto otf_cropped = [&](const float* p, const int64_t pd[3], uint64_t h) {
        int64_t win[6];
        bool whole = true;
        for (int d = 0; d < 3; ++d) {
            win[d] = pd[d] > c.d[d] ? (pd[d] - c.d[d]) / 2 : 0;
            win[d + 3] = std::min(pd[d], c.d[d]);
            whole &= win[d + 3] == pd[d];
        }
        const uint64_t key = whole ? h : key_mix(h, win, 6);
        if (const float2* hit = otf_find(*c.vp, key, win + 3)) return hit;
        std::vector<float> q;
        int64_t qd[3];
        crop_psf(p, pd, c.d, q, qd);
        return otf_build(*c.vp, key, q.data(), qd);
    };
    c.Hf = otf_cropped(psf_f, fd, hf);
    otf_pin(c.Hf, +1);
    if (c.method == 1) {
        try {
            c.Hb = otf_cropped(psf_b, bd, hb);
        } catch (...) {
            otf_pin(c.Hf, -1);
            c.Hf = nullptr;
            throw;
        }
        otf_pin(c.Hb, +1);
    }
    c.u16 = in.dtype == PK_UINT16;
    c.in_host = in.mem == PK_MEM_HOST;
    c.out_host = out.mem == PK_MEM_HOST ? out.data : nullptr;
    c.out_bytes = (c.o.save16bit ? 2 : 4) * (size_t)(c.d[0] * c.d[1] * c.d[2]);
    const VolPlan& vp = *c.vp;
    const size_t esz = c.u16 ? 2 : 4;
    const int64_t nxb = (vp.n2 * vp.n3 + vp.x.tile - 1) / vp.x.tile;
    size_t need = 2 * align_up(sizeof(float) * vp.N) + align_up(sizeof(float2) * vp.S);
    if (c.method == 0) need += 3 * align_up(sizeof(float) * vp.N) + align_up(sizeof(double) * 2 * nxb) + kAlign;
    if (c.in_host) need += align_up(esz * (size_t)(c.d[0] * c.d[1] * c.d[2]));  // staging, caller's layout
    if (c.out_host) need += align_up(c.out_bytes);
    c.raw_user = in.data;
    c.dout = c.out_host ? nullptr : out.data;
    return need;
}

// Carve the chunk's buffers from base and stage its input (copy, zero pad).
void pk_ctx_impl::chunk_place(Chunk& c, char* base, hipStream_t copy_stream) {
    const VolPlan& vp = *c.vp;
    char* p = base;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += align_up(bytes);
        return r;
    };
    c.I = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
    c.J = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
    c.A = reinterpret_cast<float2*>(take(sizeof(float2) * vp.S));
    if (c.method == 0) {
        const int64_t nxb = (vp.n2 * vp.n3 + vp.x.tile - 1) / vp.x.tile;
        c.J3 = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
        c.Y = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
        c.J4 = reinterpret_cast<float*>(take(sizeof(float) * vp.N));
        c.part = reinterpret_cast<double*>(take(sizeof(double) * 2 * nxb));
        c.lam = reinterpret_cast<float*>(take(kAlign));
    }
    const size_t esz = c.u16 ? 2 : 4;
    // the input stays in the caller's (unpadded) layout: the prep pass reads
    // it there and supplies the zero padding, the epilogue's mask reads it at
    // output positions
    const void* raw = c.raw_user;
    if (c.in_host) {
        const size_t ub = esz * (size_t)(c.d[0] * c.d[1] * c.d[2]);
        void* dst = take(ub);
        PK_HIP(hipMemcpyAsync(dst, c.raw_user, ub, hipMemcpyHostToDevice, copy_stream ? copy_stream : stream));
        raw = dst;
    }
    c.raw16 = c.u16 ? static_cast<const uint16_t*>(raw) : nullptr;
    c.rawf = c.u16 ? nullptr : static_cast<const float*>(raw);
    if (c.out_host) c.dout = take(c.out_bytes);
    c.k = 0;
    c.par = 0;
}

// Epilogue operands of a chunk (decon_lucy_function.m:177-203; no bbox or
// erosion in the step engine): output = the chunk's own (unpadded) size.
static EpiArgs chunk_epi(const Chunk& c) {
    const VolPlan& vp = *c.vp;
    EpiArgs e{};
    e.J = (c.method == 0 && c.par == 1) ? c.J3 : c.J;
    e.I = c.I;
    e.raw16 = c.raw16;
    e.rawf = c.rawf;
    e.out16 = c.o.save16bit ? static_cast<uint16_t*>(c.dout) : nullptr;
    e.outf = c.o.save16bit ? nullptr : static_cast<float*>(c.dout);
    e.n1 = vp.n1;
    e.n2 = vp.n2;
    e.o1 = c.d[0];
    e.o2 = c.d[1];
    e.o3 = c.d[2];
    e.damp = c.o.damp_factor > 1 ? (float)c.o.damp_factor : 0.f;
    e.scale = (float)c.o.scale_factor;
    e.offset = (float)c.o.decon_offset;
    e.apply_scale = (c.o.scale_factor != 1 || c.o.decon_offset != 0);
    e.use_mask = (c.o.decon_offset != 0);
    e.mask_at_out = 1;
    const size_t osz = c.o.save16bit ? 2 : 4;
    e.pair = vp.packed && c.d[0] % 2 == 0 && reinterpret_cast<uintptr_t>(c.dout) % (2 * osz) == 0;
    return e;
}

// ------------------------------------------------------------------ one step
void pk_ctx_impl::run_step(std::vector<Chunk*>& act) {
    std::vector<char> table;
    auto pos = [&](size_t bytes) {
        const size_t off = align_up(table.size());
        table.resize(off + bytes);
        return off;
    };
    auto base_x = [&](const Chunk& c, int mode) {
        XArgs a{};
        const VolPlan& vp = *c.vp;
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
        a.spec = c.A;
        return a;
    };
    auto add_x = [&](Pass<XChunk>& ps, const Chunk& c, const XArgs& a) {
        const VolPlan& vp = *c.vp;
        // chunks whose rows have a compile-time kernel are grouped by row length
        const int ctk = vp.packed ? vp.x.fp.L : -vp.x.fp.L;  // compile-time kernel key (odd rows negative)
        const int ctL = (vp.x.block.x == 256 && xrow_available(ctk, vp.x.fp.E)) ? ctk : 0;
        Group<XChunk>& g = ps[{vp.x.fp.E, a.mode, 256, ctL, (int64_t)vp.x.block.x, (int64_t)vp.x.lds}];
        g.E = vp.x.fp.E;
        g.mode = a.mode;
        g.L = ctL;
        g.chunks.push_back(XChunk{a, vp.x.fp});
        g.starts.push_back(g.starts.back() + (vp.n2 * vp.n3 + vp.x.tile - 1) / vp.x.tile);
        g.block = std::max(g.block, vp.x.block.x);
        g.lds = std::max(g.lds, vp.x.lds);
    };
    // uint16 chunks keep I as uint16 in the first half of their float I buffer
    // (I = sat(raw - bg16) holds integers below 2^16): the ratios and OMW's
    // first update read it so, with the same bits, through compile-time row
    // kernels.  Float I where other code reads it: the damp epilogue, OMW with
    // one iteration (its last update reads J = I) and no iteration.
    auto u16_est = [&](const Chunk& c) -> uint16_t* {
        const VolPlan& vp = *c.vp;
        const int ctk = vp.packed ? vp.x.fp.L : -vp.x.fp.L;
        const bool ok = c.u16 && c.numit >= (c.method == 1 ? 2 : 1) && !(c.o.damp_factor > 1) &&
                        vp.x.block.x == 256 && xrow_available(ctk, vp.x.fp.E);
        return ok ? reinterpret_cast<uint16_t*>(c.I) : nullptr;
    };
    auto add_col = [&](Pass<ColChunk>& ps, const Chunk& c, int axis, int mode, const float2* otf) {
        const VolPlan& vp = *c.vp;
        const AxisLaunch& ax = axis == 2 ? vp.y : vp.z;
        if (ax.fp.L == 1 && (mode == CM_FWD || mode == CM_INV)) return;
        ColArgs a{};
        a.data = c.A;
        a.otf = otf;
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
        const bool mixed = plan_mixed_stages(ax.fp);  // kernels of their own
        Group<ColChunk>& g = ps[{ax.fp.E, mode, ax.nt, mixed ? 1 : 0, (int64_t)ax.block.x, (int64_t)ax.lds}];
        g.E = ax.fp.E;
        g.L = mixed ? 1 : 0;
        g.mode = mode;
        g.nt = ax.nt;
        g.chunks.push_back(ColChunk{a, ax.fp, gx});
        g.starts.push_back(g.starts.back() + gx * gy);
        g.block = std::max(g.block, ax.block.x);
        g.lds = std::max(g.lds, ax.lds);
    };
    auto seal = [&](auto& ps) {
        for (auto& kv : ps) {
            auto& g = kv.second;
            using T = typename std::decay_t<decltype(g.chunks)>::value_type;
            g.dev_off = pos(sizeof(T) * g.chunks.size());
            std::memcpy(table.data() + g.dev_off, g.chunks.data(), sizeof(T) * g.chunks.size());
            g.start_off = pos(sizeof(int64_t) * g.starts.size());
            std::memcpy(table.data() + g.start_off, g.starts.data(), sizeof(int64_t) * g.starts.size());
        }
    };

    // Chunk sets, one per stream: chunks are independent, so each set runs
    // its whole pass sequence on its own stream and the launch tails of one
    // set overlap the work of the others.  Sets are balanced by voxels
    // (largest chunk first onto the lightest set); one stream when profiling.
    constexpr int kSets = kSideStreams + 1;
    const int nset = prof_on ? 1 : (int)std::min<size_t>(std::min(kSets, chunk_streams()), act.size());
    std::vector<int> set_of(act.size(), 0);
    {
        std::vector<size_t> order(act.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return act[a]->vp->N > act[b]->vp->N; });
        int64_t load[kSets] = {};
        for (size_t i : order) {
            int best = 0;
            for (int k = 1; k < nset; ++k)
                if (load[k] < load[best]) best = k;
            set_of[i] = best;
            load[best] += act[i]->vp->N;
        }
    }
    Pass<XChunk> prep[kSets], biggs_y[kSets], ratio[kSets], update[kSets];
    Pass<ColChunk> yf[kSets], zc1[kSets], yi[kSets], zc2[kSets];
    std::vector<LamChunk> lam;
    for (size_t ci = 0; ci < act.size(); ++ci) {
        Chunk& c = *act[ci];
        const int si = set_of[ci];
        if (c.k == 0) {
            XArgs a = base_x(c, c.u16 ? XM_PREP_U16 : XM_PREP_F32);
            a.raw16 = c.raw16;
            a.rawf = c.rawf;
            const VolPlan& vp = *c.vp;
            if (vp.n1 != c.d[0] || vp.n2 != c.d[1] || vp.n3 != c.d[2]) {  // padded plan: read the caller's layout
                a.s0 = c.d[0];
                a.s1 = c.d[1];
                a.s2 = c.d[2];
                a.pn2 = vp.n2;
            }
            a.bg = c.u16 ? (float)std::min(65535.0, std::max(0.0, std::round(c.o.background))) : (float)c.o.background;
            a.I = c.I;
            a.J = c.J;
            a.I16 = u16_est(c);
            if (a.I16 && c.method == 1) a.J = nullptr;  // OMW: the first update reads I (J_1 = I)
            add_x(prep[si], c, a);
            if (c.method == 0) {
                hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)std::min<int64_t>((c.vp->N + 255) / 256, 65536)), dim3(256),
                                   0, stream, c.J3, c.vp->N, 0.f);
                hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, stream, c.lam, (int64_t)1, 0.f);
                PK_HIP(hipGetLastError());
            }
        }
        if (c.k >= c.numit) continue;  // NUMIT = 0: prep only
        const int kn = c.k + 1;        // the iteration this step runs
        if (c.method == 0) {
            if (kn > 2)
                lam.push_back(LamChunk{c.part, (int)((c.vp->n2 * c.vp->n3 + c.vp->x.tile - 1) / c.vp->x.tile), c.lam});
            XArgs a = base_x(c, XM_BIGGS_Y);
            a.J = c.par ? c.J3 : c.J;
            a.J3 = c.par ? c.J : c.J3;
            a.Y = c.Y;
            a.lambda = c.lam;
            add_x(biggs_y[si], c, a);
        }
        const int cax = c.vp->conv_axis, oax = cax == 3 ? 2 : 3;  // pass order of the chunk's plan
        add_col(yf[si], c, oax, CM_FWD, nullptr);
        add_col(zc1[si], c, cax, CM_CONV, c.Hf);
        add_col(yi[si], c, oax, CM_INV, nullptr);
        {
            uint16_t* I16 = u16_est(c);
            XArgs a = base_x(c, I16 ? XM_RATIO_U16 : XM_RATIO);
            a.I = c.I;
            a.I16 = I16;
            add_x(ratio[si], c, a);
        }
        add_col(zc2[si], c, cax, c.method == 1 ? CM_CONV : CM_CONVJ, c.method == 1 ? c.Hb : c.Hf);
        const bool last = kn == c.numit;  // the last update writes the output through the epilogue
        if (c.method == 1) {
            uint16_t* I16 = kn == 1 ? u16_est(c) : nullptr;  // J_1 = I, read as uint16
            XArgs a = base_x(c, last ? XM_OMW_FINAL : I16 ? XM_OMW_UPDATE_U16 : XM_OMW_UPDATE);
            a.J = c.J;
            a.I16 = I16;
            a.r2c_out = !last;
            if (last) a.ep = chunk_epi(c);
            add_x(update[si], c, a);
        } else {
            XArgs a = base_x(c, last ? XM_BIGGS_FINAL : XM_BIGGS_UPDATE);
            a.Y = c.Y;
            a.Jn = c.par ? c.J : c.J3;  // x_{k-2}'s buffer receives x_k
            a.J4 = c.J4;
            a.acc = (kn >= 2 && !last);
            a.write_j4 = !last;
            a.partials = c.part;
            if (last) a.ep = chunk_epi(c);
            add_x(update[si], c, a);
        }
    }
    for (int k = 0; k < nset; ++k) {
        seal(prep[k]);
        seal(biggs_y[k]);
        seal(ratio[k]);
        seal(update[k]);
        seal(yf[k]);
        seal(zc1[k]);
        seal(yi[k]);
        seal(zc2[k]);
    }
    size_t lam_off = 0;
    if (!lam.empty()) {
        lam_off = pos(sizeof(LamChunk) * lam.size());
        std::memcpy(table.data() + lam_off, lam.data(), sizeof(LamChunk) * lam.size());
    }
    if (table.empty()) return;
    const int slot = ws.tab_slot;
    ws.tab_slot ^= 1;
    if (ws.tab_ev[slot]) PK_HIP(hipEventSynchronize(ws.tab_ev[slot]));  // its previous upload has been read
    else PK_HIP(hipEventCreateWithFlags(&ws.tab_ev[slot], hipEventDisableTiming));
    HostBuf& hb = ws.tab_host[slot];
    if (hb.bytes < table.size()) {
        if (hb.p) PK_HIP(hipHostFree(hb.p));
        hb.p = nullptr;
        hb.bytes = 0;
        const size_t want = std::max<size_t>(table.size() * 2, 64 << 10);
        PK_HIP(hipHostMalloc(&hb.p, want, hipHostMallocDefault));
        hb.bytes = want;
    }
    std::memcpy(hb.p, table.data(), table.size());
    // the device table is reused two steps later, after this step's kernels in stream order
    ensure(ws.tab_dev[slot], table.size());
    PK_HIP(hipMemcpyAsync(ws.tab_dev[slot].p, hb.p, table.size(), hipMemcpyHostToDevice, stream));
    PK_HIP(hipEventRecord(ws.tab_ev[slot], stream));
    char* tab = ws.tab_dev[slot].as<char>();

    auto run_x = [&](const Pass<XChunk>& ps, const char* name, hipStream_t st) {
        for (auto& kv : ps) {
            const auto& g = kv.second;
            hipEvent_t e0 = prof_mark();
            const XChunk* chd = reinterpret_cast<const XChunk*>(tab + g.dev_off);
            const int64_t* std_ = reinterpret_cast<const int64_t*>(tab + g.start_off);
            if (g.L && launch_xrow_batch(g.L, g.E, dim3((unsigned)g.starts.back()), g.lds, st, g.mode, chd, std_,
                                         (int)g.chunks.size()))
                PK_HIP(hipGetLastError());
            else
                PK_HIP(launch_xpass_batch(g.E, dim3((unsigned)g.starts.back()), dim3(g.block), g.lds, st, g.mode,
                                          chd, std_, (int)g.chunks.size()));
            prof_done(name, 0, e0);
        }
    };
    auto run_c = [&](const Pass<ColChunk>& ps, const char* name, hipStream_t st) {
        for (auto& kv : ps) {
            const auto& g = kv.second;
            hipEvent_t e0 = prof_mark();
            PK_HIP(launch_colpass_batch(g.E, g.nt, g.L != 0, dim3((unsigned)g.starts.back()), dim3(g.block), g.lds,
                                        st, g.mode, reinterpret_cast<const ColChunk*>(tab + g.dev_off),
                                        reinterpret_cast<const int64_t*>(tab + g.start_off), (int)g.chunks.size()));
            prof_done(name, 0, e0);
        }
    };
    if (!lam.empty()) {
        hipLaunchKernelGGL(k_lambda_batch, dim3((unsigned)lam.size()), dim3(256), 0, stream,
                           reinterpret_cast<const LamChunk*>(tab + lam_off), kEpsF);
        PK_HIP(hipGetLastError());
    }
    // fork: the side streams start after the table upload and the step lengths
    if (nset > 1) {
        if (!ev_fork) PK_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        PK_HIP(hipEventRecord(ev_fork, stream));
        for (int k = 1; k < nset; ++k) {
            if (!side[k - 1]) PK_HIP(hipStreamCreateWithFlags(&side[k - 1], hipStreamNonBlocking));
            PK_HIP(hipStreamWaitEvent(side[k - 1], ev_fork, 0));
        }
    }
    for (int k = 0; k < nset; ++k) {
        hipStream_t st = k == 0 ? stream : side[k - 1];
        run_x(prep[k], "b_x_prep", st);
        run_x(biggs_y[k], "b_x_biggs_y", st);
        run_c(yf[k], "b_y_fwd", st);
        run_c(zc1[k], "b_z_conv", st);
        run_c(yi[k], "b_y_inv", st);
        run_x(ratio[k], "b_x_ratio", st);
        run_c(yf[k], "b_y_fwd", st);
        run_c(zc2[k], "b_z_conv", st);
        run_c(yi[k], "b_y_inv", st);
        run_x(update[k], "b_x_update", st);
    }
    // join: later work on the engine's stream sees every set's results
    for (int k = 1; k < nset; ++k) {
        if (!ev_join[k - 1]) PK_HIP(hipEventCreateWithFlags(&ev_join[k - 1], hipEventDisableTiming));
        PK_HIP(hipEventRecord(ev_join[k - 1], side[k - 1]));
        PK_HIP(hipStreamWaitEvent(stream, ev_join[k - 1], 0));
    }
    for (Chunk* cp : act) {
        Chunk& c = *cp;
        if (c.k < c.numit) {
            c.k++;
            if (c.method == 0) c.par ^= 1;
        } else if (c.numit == 0) {
            c.k = 1;  // prep done
        }
    }
}

bool Chunk::done() const { return k >= std::max(numit, 1); }

// Epilogue (decon_lucy_function.m:177-197) and output copy of a finished chunk.
void pk_ctx_impl::chunk_finish(Chunk& c, hipStream_t copy_stream) {
    // NUMIT > 0: the last update wrote the output (fused epilogue)
    if (c.numit == 0) {
        const EpiArgs e = chunk_epi(c);
        hipLaunchKernelGGL(k_epilogue, dim3(epi_grid(e)), dim3(256), 0, stream, e);
        PK_HIP(hipGetLastError());
    }
    if (c.out_host) {
        if (copy_stream) copy_after_stream(copy_stream);  // after the chunk's last step (and epilogue)
        PK_HIP(hipMemcpyAsync(c.out_host, c.dout, c.out_bytes, hipMemcpyDeviceToHost, copy_stream ? copy_stream : stream));
    }
}

void pk_ctx_impl::copy_after_stream(hipStream_t copy_stream) {
    if (!ev_copy) PK_HIP(hipEventCreateWithFlags(&ev_copy, hipEventDisableTiming));
    PK_HIP(hipEventRecord(ev_copy, stream));
    PK_HIP(hipStreamWaitEvent(copy_stream, ev_copy, 0));
}

// ------------------------------------------------------------------ batch
void pk_ctx_impl::decon_batch(int method, const pk_batch_item* items, int n, const float* psf_f, const int64_t fd[3],
                              const float* psf_b, const int64_t bd[3], int numit, const pk_opts* o) {
    if (!items || n < 1 || !o || !psf_f || numit < 0) fail(PK_ERR_INVALID_ARG, "empty batch or null argument");
    if (method == 1 && !psf_b) fail(PK_ERR_INVALID_ARG, "PSF_b missing");
    for (int i = 0; i < n; ++i) chunk_validate(items[i].in, items[i].out, *o);
    const bool u16 = items[0].in.dtype == PK_UINT16;
    for (int i = 1; i < n; ++i)
        if ((items[i].in.dtype == PK_UINT16) != u16) fail(PK_ERR_INVALID_ARG, "all chunks must share one input class");
    std::map<std::vector<int64_t>, int> distinct;
    for (int i = 0; i < n; ++i)
        distinct[{smooth_up(items[i].in.dims[0]), smooth_up(items[i].in.dims[1]), smooth_up(items[i].in.dims[2])}]++;
    otf_max_entries = std::max<size_t>(otf_max_entries, 2 * distinct.size() + 4);
    // one hash per PSF for the whole batch
    const uint64_t hf = psf_key(psf_f, fd);
    const uint64_t hb = method == 1 ? psf_key(psf_b, bd) : 0;
    std::vector<Chunk> ch((size_t)n);
    std::vector<size_t> offs((size_t)n);
    size_t arena = 0;
    struct Unpin {  // release the OTF pins however the call ends
        pk_ctx_impl* c;
        std::vector<Chunk>* ch;
        ~Unpin() {
            for (auto& k : *ch) {
                if (k.Hf) c->otf_pin(k.Hf, -1);
                if (k.Hb) c->otf_pin(k.Hb, -1);
            }
        }
    } unpin{this, &ch};
    for (int i = 0; i < n; ++i) {
        ch[(size_t)i].method = method;
        ch[(size_t)i].numit = numit;
        ch[(size_t)i].o = *o;
        offs[(size_t)i] = arena;
        arena += chunk_prepare(ch[(size_t)i], items[i].in, items[i].out, psf_f, fd, hf, psf_b, bd, hb);
    }
    ensure(ws.batch, arena);
    std::vector<Chunk*> act;
    for (int i = 0; i < n; ++i) {
        chunk_place(ch[(size_t)i], ws.batch.as<char>() + offs[(size_t)i]);
        act.push_back(&ch[(size_t)i]);
    }
    for (int s = 0; s < std::max(numit, 1); ++s) run_step(act);
    for (auto& c : ch) chunk_finish(c);
    PK_HIP(hipStreamSynchronize(stream));
    prof_flush();
}

}  // namespace pk
