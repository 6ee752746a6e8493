// pk_engine.hip -- host side of the gfx950 deconvolution engine.
//
// Iteration schedules (one launch per line; x = dim-1 row pass, y / z =
// strided passes; every elementwise step of the reference rides inside the
// first or last pass of a transform):
//
//   OMW (decon_lucy_omw_function.m:105-112), per iteration:
//     [x PREP or previous x UPDATE]  J -> A = R2C_x(J)
//     y FWD;  z FWD*OTF_f*INV;  y INV
//     x RATIO      c = C2R_x(A); r = I ./ max(c, eps); A = R2C_x(r)
//     y FWD;  z FWD*OTF_b*INV;  y INV
//     x UPDATE     t = C2R_x(A); J = max(t .* J, 0); A = R2C_x(J) unless last
//
//   Biggs RL (decon_lucy_function.m:121-143), per iteration k:
//     [k >= 3] lambda from the block partials of the previous update
//     x BIGGS_Y    Y = max(J2 + lambda (J2 - J3), 0); A = R2C_x(Y)
//     y FWD;  z FWD*H*INV;  y INV
//     x RATIO
//     y FWD;  z FWD*conj(H)*INV;  y INV
//     x BIGGS_UPDATE  Jn = max(Y .* C2R_x(A), 0) (into J3's buffer);
//                     J4 = Jn - Y, with the two dot products of the next
//                     step length accumulated before J4 is overwritten
#include "pk_engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "pk_launch.hpp"
#include "pk_misc.hpp"

namespace pk {

static const float kEps = 2.220446049250313080847e-16f;  // MATLAB eps (2^-52), exact in single

void fail(pk_status st, const std::string& msg) { throw Error{st, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        pk_status st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PK_ERR_NO_MEMORY : PK_ERR_HIP;
        fail(st, std::string(what) + ": " + hipGetErrorString(e));
    }
}

bool smooth235(int64_t n) {
    if (n < 1) return false;
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

// ------------------------------------------------------------------ DevBuf
DevBuf::~DevBuf() { release(); }
void DevBuf::release() {
    if (p) {
        if (pooled && owner) owner->block_put(p, bytes);
        else (void)hipFree(p);
        if (owner) owner->live_bytes -= bytes;
    }
    p = nullptr;
    bytes = 0;
    pooled = false;
}

// Size class of a cached block: the size rounded up to 1/8 of its highest
// power of two (at most 12.5 % slack), 1 MiB at least.
static size_t block_class(size_t b) {
    if (b <= (size_t(1) << 20)) return size_t(1) << 20;
    size_t p = size_t(1) << (63 - __builtin_clzll((unsigned long long)b));
    const size_t step = p / 8;
    return (b + step - 1) / step * step;
}

void pk_ctx_impl::block_put(void* p, size_t bytes) {
    block_cache.emplace(bytes, p);
    block_cache_bytes += bytes;
}

void pk_ctx_impl::block_cache_clear() {
    if (block_cache.empty()) return;
    for (auto& kv : block_cache) (void)hipFree(kv.second);
    block_cache.clear();
    block_cache_bytes = 0;
}

void pk_ctx_impl::alloc_pooled(DevBuf& b, size_t bytes) {
    b.release();
    if (bytes == 0) return;
    const size_t cls = block_class(bytes);
    auto it = block_cache.find(cls);
    if (it == block_cache.end()) {
        if (budget && live_bytes + block_cache_bytes + cls > budget) block_cache_clear();
        alloc(b, cls);  // budget check and hipMalloc
        b.pooled = true;
        return;
    }
    b.p = it->second;
    b.bytes = cls;
    b.owner = this;
    b.pooled = true;
    block_cache.erase(it);
    block_cache_bytes -= cls;
    live_bytes += cls;
}

void pk_ctx_impl::alloc(DevBuf& b, size_t bytes) {
    b.release();
    if (bytes == 0) return;
    if (budget && live_bytes + block_cache_bytes + bytes > budget) block_cache_clear();
    if (budget && live_bytes + bytes > budget)
        fail(PK_ERR_BUDGET, "device allocation of " + std::to_string(bytes) + " bytes exceeds the HBM budget (" +
                                std::to_string(live_bytes) + " of " + std::to_string(budget) + " bytes in use)");
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(PK_ERR_NO_MEMORY, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
    }
    b.p = p;
    b.bytes = bytes;
    b.owner = this;
    live_bytes += bytes;
}

void pk_ctx_impl::ensure(DevBuf& b, size_t bytes) {
    if (b.bytes < bytes) alloc(b, bytes);
}

// ------------------------------------------------------------------ profiling
hipEvent_t pk_ctx_impl::prof_mark() {
    if (!prof_on) return nullptr;
    hipEvent_t e;
    if (!ev_pool.empty()) {
        e = ev_pool.back();
        ev_pool.pop_back();
    } else {
        PK_HIP(hipEventCreate(&e));
    }
    PK_HIP(hipEventRecord(e, stream));
    return e;
}

void pk_ctx_impl::prof_done(const char* name, double bytes, hipEvent_t e0) {
    if (!prof_on || !e0) return;
    hipEvent_t e1 = prof_mark();
    prof_pending.push_back(ProfPending{name, bytes, e0, e1});
}

void pk_ctx_impl::prof_flush() {
    for (auto& p : prof_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
            ProfAgg& a = prof[p.name];
            a.launches++;
            a.total_ms += ms;
            a.bytes = p.bytes;
        }
        ev_pool.push_back(p.e0);
        ev_pool.push_back(p.e1);
    }
    prof_pending.clear();
}

// ------------------------------------------------------------------ planning
static const int kEs[] = {1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20, 24, 25, 27, 30, 32};

// Launch-shape overrides for tuning sweeps, read once per process:
// PK_FFT_EX / PK_FFT_EY / PK_FFT_EZ force the values-per-thread count of an
// axis when it is a feasible plan for that length; PK_FFT_TY / PK_FFT_TZ cap
// the column-tile width of the strided axes.  Unset (the default) means the
// built-in choice.
struct Tune {
    int e[3] = {0, 0, 0};
    int t[3] = {0, 0, 0};
    bool generic_x = false;  // PK_FFT_GENERIC_X=1: always the runtime-plan row kernel
    bool graphs = true;      // PK_GRAPHS=0: launch the iteration loops kernel by kernel
    // XCD-aware tile order in the strided passes (PK_XCD_REMAP=0 turns it off);
    // measured: 512x512x256 z_conv 0.190 -> 0.174 ms, 2048-point dim-2 passes
    // 5.19 -> 4.64 ms
    int xcd = 1;
    int slab = -1;           // PK_SLAB_PLANES: planes per slab (0 = off), -1 = built-in
    // PK_CHUNK_STREAMS: streams of a batched step (1..8).  Measured on 64 chunks
    // (OMW-2 / Biggs-10 Gvoxel/s, one box, runs interleaved, twelve hardware
    // queues): 1 -> 16.65 / 3.08, 2 -> 17.97 / 3.35, 4 -> 18.35-18.39 / 3.42,
    // 6 -> 17.78 / 3.34, 8 -> 17.92-18.06 / 3.34-3.35; the served mix at 4
    // streams had 5.6 % more capacity than at 8
    int streams = 4;
    // PK_TUNE_COL_CANDS: strided-axis shapes the autotuner times per axis, best-ranked first
    // (64 chunks, OMW-2 Gvoxel/s: 4 -> 17.5, 8 -> 17.0)
    int col_cands = 4;
    Tune() {
        if (const char* v = std::getenv("PK_CHUNK_STREAMS")) streams = std::max(1, std::min(8, std::atoi(v)));
        if (const char* v = std::getenv("PK_TUNE_COL_CANDS")) col_cands = std::max(1, std::min(16, std::atoi(v)));
        if (const char* v = std::getenv("PK_SLAB_PLANES")) slab = std::atoi(v);
        if (const char* v = std::getenv("PK_GRAPHS")) graphs = std::atoi(v) != 0;
        if (const char* v = std::getenv("PK_XCD_REMAP")) xcd = std::atoi(v);
        if (const char* v = std::getenv("PK_FFT_GENERIC_X")) generic_x = std::atoi(v) != 0;
        const char* names_e[3] = {"PK_FFT_EX", "PK_FFT_EY", "PK_FFT_EZ"};
        const char* names_t[3] = {nullptr, "PK_FFT_TY", "PK_FFT_TZ"};
        for (int i = 0; i < 3; ++i) {
            if (const char* v = std::getenv(names_e[i])) e[i] = std::atoi(v);
            if (names_t[i])
                if (const char* v = std::getenv(names_t[i])) t[i] = std::atoi(v);
        }
    }
};
static const Tune& tune() {
    static const Tune t;
    return t;
}

// Factor L into radices that all divide E; E is a multiple of every prime of L.
// score = stage_w * stages + e_w * |E - target_e|, lowest wins.
static FftPlan plan_1d(int L, int target_e, int max_threads, int force_e, int stage_w, int e_w) {
    FftPlan best{};
    best.L = L;
    if (L == 1) {
        best.P = 1;
        best.E = 1;
        best.nst = 0;
        return best;
    }
    int best_score = 1 << 30;
    bool found = false;
    for (int E : kEs) {
        if (L % E) continue;
        if (L / E > max_threads) continue;
        bool ok = true;
        for (int p : {2, 3, 5})
            if (L % p == 0 && E % p != 0) ok = false;
        if (!ok || !e_supported(E)) continue;
        FftPlan c{};
        c.L = L;
        c.E = E;
        c.P = L / E;
        int rem = L, ns = 1;
        while (rem > 1) {
            const int r = radix_for(rem, E);
            if (!r || c.nst >= kMaxStages) {
                ok = false;
                break;
            }
            c.R[c.nst] = r;
            c.Ns[c.nst] = ns;
            c.mag[c.nst] = ns > 1 ? (uint32_t)((1ull << 32) / (uint64_t)ns + 1ull) : 0u;
            c.tws[c.nst] = L / (ns * r);
            c.nst++;
            ns *= r;
            rem /= r;
        }
        if (!ok) continue;
        const int score = (force_e == E) ? -1 : c.nst * stage_w + e_w * std::abs(E - target_e);
        if (score < best_score) {
            best_score = score;
            best = c;
            found = true;
        }
    }
    if (!found) fail(PK_ERR_UNSUPPORTED_SIZE, "no transform plan for length " + std::to_string(L));
    return best;
}

static void fill_twiddles(std::vector<float2>& t, int L) {
    t.resize(L);
    for (int e = 0; e < L; ++e) {
        // exact at quarter turns; long double elsewhere, rounded once
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)e / (long double)L;
        float c = (float)cosl(a), s = (float)(-sinl(a));
        if (4 * e == L) { c = 0.f; s = -1.f; }
        if (2 * e == L) { c = -1.f; s = 0.f; }
        if (4 * e == 3 * L) { c = 0.f; s = 1.f; }
        t[e] = make_float2(c, s);
    }
}

// Row-pass launch shapes for length Lx (complex points), best first by the
// built-in rule: with a compile-time kernel for the length, E near 8
// (measured at 512x512x256, ratio pass: E = 4 / 8 / 16 -> 0.209 / 0.175 /
// 0.170 ms); the runtime-plan kernel spends more on index arithmetic and
// prefers few values per thread (E = 4 / 8 / 16 -> 0.24 / 0.26 / 0.29 ms).
static AxisLaunch x_axis(const FftPlan& fp, bool compiled = false) {
    AxisLaunch ax;
    ax.fp = fp;
    ax.lpad = row_pitch(fp.L);
    if (compiled) {
        // compile-time kernel: whole 256-thread blocks, rows packed per wavefront
        ax.tile = xrow_tile(fp.P);
        ax.block = dim3(kMaxThreads);
        ax.lds = std::max<size_t>((size_t)xrow_lds_rows(fp.P) * ax.lpad * sizeof(float2),
                                  2 * sizeof(double) * kMaxThreads);
        return ax;
    }
    int rt = std::max(1, kMaxThreads / fp.P);
    while (rt > 1 && (size_t)rt * ax.lpad * sizeof(float2) > 65536) rt--;
    ax.tile = rt;
    ax.block = dim3(rt * fp.P);
    ax.lds = std::max<size_t>((size_t)rt * ax.lpad * sizeof(float2), 2 * sizeof(double) * kMaxThreads);
    return ax;
}
// Plan of a compile-time row kernel: L and E are all the kernel takes from it
// (its stages are compile-time constants; the table pointer is set by the caller).
static FftPlan ct_row_plan(int L, int E) {
    FftPlan fp{};
    fp.L = L;
    fp.E = E;
    fp.P = L / E;
    fp.nst = ct_nst(L, E);
    return fp;
}
static std::vector<AxisLaunch> x_candidates(int Lx, bool packed) {
    std::vector<AxisLaunch> out;
    if (tune().e[0] > 0) {
        out.push_back(x_axis(plan_1d(Lx, tune().e[0], kMaxThreads, tune().e[0], 10, 4)));
        return out;
    }
    // every feasible values-per-thread count, scored toward 8 (compile-time
    // kernels) and toward 4 (runtime-plan kernel)
    std::vector<std::pair<int, FftPlan>> ct, gen;
    for (int E : kEs) {
        if (Lx % E || Lx / E > kMaxThreads) continue;
        if (!tune().generic_x && xrow_available(packed ? Lx : -Lx, E)) {
            const FftPlan fp = ct_row_plan(Lx, E);
            ct.emplace_back(fp.nst * 10 + 4 * std::abs(E - 8), fp);
        }
        bool ok = e_supported(E);
        for (int pr : {2, 3, 5})
            if (Lx % pr == 0 && E % pr != 0) ok = false;
        if (!ok) continue;
        const FftPlan fp = plan_1d(Lx, E, kMaxThreads, E, 10, 4);
        if (fp.E != E) continue;
        gen.emplace_back(fp.nst * 10 + 4 * std::abs(E - 4), fp);
    }
    if (ct.empty() && gen.empty()) fail(PK_ERR_UNSUPPORTED_SIZE, "no transform plan for length " + std::to_string(Lx));
    auto by = [](const auto& a, const auto& b) { return a.first < b.first; };
    std::stable_sort(ct.begin(), ct.end(), by);
    std::stable_sort(gen.begin(), gen.end(), by);
    for (auto& c : ct) out.push_back(x_axis(c.second, true));
    for (auto& g : gen) {
        if (out.size() >= 4) break;
        bool dup = false;
        for (auto& o : out) dup |= o.fp.E == g.second.E;
        if (!dup) out.push_back(x_axis(g.second));
    }
    if (out.size() > 4) out.resize(4);
    return out;
}

// Plan of values-per-thread E for a length E does not cover (a prime of L
// missing from E): a radix dividing E where one divides the remaining length,
// else the largest of kMixedRadices dividing it (that stage runs from LDS).
// nst = 0 when no such plan exists; *mm receives the number of LDS-fed stages.
static FftPlan plan_mixed(int L, int E, int* mm) {
    FftPlan c{};
    c.L = L;
    c.E = E;
    c.P = L / E;
    *mm = 0;
    int rem = L, ns = 1;
    while (rem > 1) {
        int r = radix_for(rem, E);
        if (!r) {
            for (int q : kMixedRadices)
                if (rem % q == 0) {
                    r = q;
                    break;
                }
            ++*mm;
        }
        if (!r || c.nst >= kMaxStages) {
            c.nst = 0;
            return c;
        }
        c.R[c.nst] = r;
        c.Ns[c.nst] = ns;
        c.mag[c.nst] = ns > 1 ? (uint32_t)((1ull << 32) / (uint64_t)ns + 1ull) : 0u;
        c.tws[c.nst] = L / (ns * r);
        c.nst++;
        ns *= r;
        rem /= r;
    }
    return c;
}
// Column plan for (L, E): the all-register plan when E holds every prime of
// L, else a mixed plan with at most two LDS-fed stages; E = 0 when neither.
static FftPlan col_plan(int L, int E, int* mm) {
    *mm = 0;
    bool full = true;
    for (int pr : {2, 3, 5})
        if (L % pr == 0 && E % pr != 0) full = false;
    if (full) {
        const FftPlan fp = plan_1d(L, E, kMaxThreads * 4, E, 100, 1);
        if (fp.E == E) return fp;
    }
    FftPlan fp = plan_mixed(L, E, mm);
    if (L == 1 || fp.nst == 0 || *mm > 2 || !col_mixed_available(E, kMaxThreads)) fp.E = 0;
    return fp;
}

// Strided-pass (y, z) launch shapes for length L: tiles of T adjacent
// columns (T float2 per row segment), sorted by the built-in rule: T = 16
// (128-byte segments) first, then fewest stages, then E near 16, then the
// smallest block.  Long columns use 512- / 1024-thread blocks and up to
// 160 KB of LDS (T * L * 8 bytes) to keep T >= 8.
static std::vector<AxisLaunch> col_candidates(int L, int axis) {
    std::vector<std::pair<int, AxisLaunch>> c;
    for (int E : kEs) {
        if (L % E) continue;
        if (tune().e[axis] > 0 && E != tune().e[axis] && L != 1) continue;
        if (!e_supported(E)) continue;
        int mm = 0;
        const FftPlan fp = L == 1 ? plan_1d(L, E, kMaxThreads * 4, E, 100, 1) : col_plan(L, E, &mm);
        if (fp.E != E) continue;
        for (int nt : {kMaxThreads, 512, 1024}) {
            if (nt != kMaxThreads && (mm > 0 || !col_wide_available(E, nt))) continue;
            if (fp.P > nt) continue;
            int T = 1;
            while (T * 2 <= 16 && T * 2 * fp.P <= nt && (size_t)T * 2 * L * sizeof(float2) <= kMaxLds) T *= 2;
            if (tune().t[axis] > 0) T = std::min(T, tune().t[axis]);
            // also 32-column tiles (256-byte row segments) where they fit: ranked
            // after the 16-column shape of the same E, for the autotuner to measure
            for (int Tc : {T, 2 * T}) {
                if (Tc > T && (T != 16 || Tc * fp.P > nt || (size_t)Tc * L * sizeof(float2) > kMaxLds ||
                               tune().t[axis] > 0))
                    continue;
                AxisLaunch ax;
                ax.fp = fp;
                ax.tile = Tc;
                ax.nt = nt;
                ax.block = dim3(Tc * fp.P);
                ax.lds = std::max<size_t>((size_t)Tc * L * sizeof(float2), 16);
                c.emplace_back((16 - T) * 1000 + (fp.nst + mm) * 100 + std::abs(E - 16) * 2 +
                                   (nt > kMaxThreads ? 1 : 0) + (Tc > T ? 50 : 0),
                               ax);
            }
        }
    }
    if (c.empty()) fail(PK_ERR_UNSUPPORTED_SIZE, "no transform plan for length " + std::to_string(L));
    std::stable_sort(c.begin(), c.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<AxisLaunch> out;
    for (auto& e : c) {  // one block size per (E, tile): the smallest that fits ranks first
        bool dup = false;
        for (auto& o : out) dup |= o.fp.E == e.second.fp.E && o.tile == e.second.tile;
        if (!dup) out.push_back(e.second);
    }
    return out;
}

// A column launch with given values per thread, block size and tile width;
// false when the combination is not available for length L.
static bool col_axis(int L, int E, int nt, int T, AxisLaunch& ax) {
    if (L % E || !e_supported(E) || T < 1) return false;
    if (nt != kMaxThreads && !col_wide_available(E, nt)) return false;
    int mm = 0;
    const FftPlan fp = L == 1 ? plan_1d(L, E, kMaxThreads * 4, E, 100, 1) : col_plan(L, E, &mm);
    if (mm > 0 && nt != kMaxThreads) return false;  // LDS-fed stages: 256-thread kernels only
    if (fp.E != E || fp.P * T > nt || (size_t)T * L * sizeof(float2) > kMaxLds) return false;
    ax = AxisLaunch{};
    ax.fp = fp;
    ax.tile = T;
    ax.nt = nt;
    ax.block = dim3(T * fp.P);
    ax.lds = std::max<size_t>((size_t)T * L * sizeof(float2), 16);
    return true;
}

// Tuned shapes per padded size, shared by the process's contexts and kept in
// the text file PK_TUNE_FILE when set (one line per size:
// n1 n2 n3  xE  yE yNT yT  zE zNT zT  conv_axis).
struct Wisdom {
    std::mutex mu;
    std::map<std::vector<int64_t>, std::array<int, 8>> table;
    bool loaded = false;
    static Wisdom& get() {
        static Wisdom w;
        return w;
    }
    void load() {
        if (loaded) return;
        loaded = true;
        const char* f = std::getenv("PK_TUNE_FILE");
        if (!f) return;
        FILE* fp = std::fopen(f, "r");
        if (!fp) return;
        long long a, b, c;
        int v[8];
        while (std::fscanf(fp, "%lld %lld %lld %d %d %d %d %d %d %d %d", &a, &b, &c, &v[0], &v[1], &v[2], &v[3], &v[4],
                           &v[5], &v[6], &v[7]) == 11)
            table[{a, b, c}] = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
        std::fclose(fp);
    }
    void store(const std::vector<int64_t>& k, const std::array<int, 8>& v) {
        table[k] = v;
        const char* f = std::getenv("PK_TUNE_FILE");
        if (!f) return;
        FILE* fp = std::fopen(f, "a");
        if (!fp) return;
        std::fprintf(fp, "%lld %lld %lld %d %d %d %d %d %d %d %d\n", (long long)k[0], (long long)k[1], (long long)k[2],
                     v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
        std::fclose(fp);
    }
};

// Launch shapes of one wisdom line {xE, yE, yNT, yT, zE, zNT, zT, conv_axis}
// applied to a plan; false (plan unchanged) when a shape is infeasible.
bool pk_ctx_impl::apply_shapes(VolPlan& vp, const std::array<int, 8>& v) {
    const int Lx = (int)(vp.packed ? vp.n1 / 2 : vp.n1);
    const bool ctx =
        !tune().generic_x && v[0] > 0 && Lx % v[0] == 0 && xrow_available(vp.packed ? Lx : -Lx, v[0]);
    if (!ctx && (v[0] < 1 || Lx % v[0] != 0)) return false;
    const FftPlan fx = ctx ? ct_row_plan(Lx, v[0]) : plan_1d(Lx, v[0], kMaxThreads, v[0], 10, 4);
    AxisLaunch ay, az;
    if (fx.E != v[0] || !col_axis((int)vp.n2, v[1], v[2], v[3], ay) || !col_axis((int)vp.n3, v[4], v[5], v[6], az) ||
        !(v[7] == 2 || v[7] == 3))
        return false;
    const float2 *tx = vp.x.fp.tw, *ty = vp.y.fp.tw, *tz = vp.z.fp.tw;  // the plan's tables, per length
    vp.x = x_axis(fx, ctx);
    vp.x.fp.tw = tx;
    vp.y = ay;
    vp.y.fp.tw = ty;
    vp.z = az;
    vp.z.fp.tw = tz;
    vp.conv_axis = v[7];
    vp.tuned = true;
    return true;
}

void pk_ctx_impl::set_shapes(int64_t n1, int64_t n2, int64_t n3, const std::array<int, 8>& v) {
    const std::vector<int64_t> key{n1, n2, n3};
    VolPlan vp = plan(n1, n2, n3);
    if (!apply_shapes(vp, v)) fail(PK_ERR_INVALID_ARG, "launch shapes not feasible for this size");
    plans[key] = vp;
    clear_graphs();  // captured loops hold the old launch shapes
    Wisdom& w = Wisdom::get();
    std::lock_guard<std::mutex> lk(w.mu);
    w.load();
    w.table[key] = v;
}

// W_{n1}^k for k in [0, n1/2] (the split step of packed real rows), exact at
// quarter and half turns, long double elsewhere, rounded once.
static void fill_real_twiddles(std::vector<float2>& h, int64_t n1) {
    const int M = (int)(n1 / 2);
    h.resize(M + 1);
    for (int k = 0; k <= M; ++k) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n1;
        float c = (float)cosl(a), s = (float)(-sinl(a));
        if (4 * k == n1) { c = 0.f; s = -1.f; }
        if (2 * k == n1) { c = -1.f; s = 0.f; }
        h[k] = make_float2(c, s);
    }
}

VolPlan pk_ctx_impl::plan_shapes(int64_t n1, int64_t n2, int64_t n3) const {
    for (int64_t n : {n1, n2, n3})
        if (!smooth235(n))
            fail(PK_ERR_UNSUPPORTED_SIZE, "axis length " + std::to_string(n) +
                                              " is not 2/3/5-smooth; pad it with findGoodFactorNumber(n, 1, false, 5)");
    if (n1 > 16384 || n2 > 8192 || n3 > 8192) fail(PK_ERR_UNSUPPORTED_SIZE, "axis longer than the supported maximum");
    VolPlan vp;
    vp.n1 = n1;
    vp.n2 = n2;
    vp.n3 = n3;
    vp.h1 = n1 / 2 + 1;
    vp.hp = (vp.h1 + 15) / 16 * 16;
    vp.N = n1 * n2 * n3;
    vp.S = vp.hp * n2 * n3;
    vp.Sa = vp.h1 * n2 * n3;
    vp.packed = (n1 % 2 == 0);
    // x (rows), y, z (columns): the first candidate is the built-in choice
    vp.x = x_candidates(vp.row_len(), vp.packed)[0];
    vp.y = col_candidates((int)n2, 1)[0];
    vp.z = col_candidates((int)n3, 2)[0];
    {  // shapes measured earlier for this size (this process or PK_TUNE_FILE)
        Wisdom& w = Wisdom::get();
        std::lock_guard<std::mutex> lk(w.mu);
        w.load();
        auto t = w.table.find({n1, n2, n3});
        if (t != w.table.end()) {
            VolPlan tuned = vp;
            if (const_cast<pk_ctx_impl*>(this)->apply_shapes(tuned, t->second)) vp = tuned;
        }
    }
    return vp;
}

void twiddle_host_tables(const VolPlan& vp, std::vector<float2>& tx, std::vector<float2>& ty,
                         std::vector<float2>& tz, std::vector<float2>& rt) {
    fill_twiddles(tx, vp.row_len());
    fill_twiddles(ty, (int)vp.n2);
    fill_twiddles(tz, (int)vp.n3);
    rt.clear();
    if (vp.packed) fill_real_twiddles(rt, vp.n1);
}

const VolPlan& pk_ctx_impl::plan(int64_t n1, int64_t n2, int64_t n3) {
    std::vector<int64_t> key{n1, n2, n3};
    auto it = plans.find(key);
    if (it != plans.end()) return it->second;
    VolPlan vp = plan_shapes(n1, n2, n3);
    // device twiddle tables, one per transform length, shared by the plans
    auto table = [&](int L) -> const float2* {
        auto t = tw.find(L);
        if (t != tw.end()) return t->second->as<float2>();
        std::vector<float2> h;
        fill_twiddles(h, L);
        auto b = std::make_unique<DevBuf>();
        alloc(*b, sizeof(float2) * (size_t)L);
        PK_HIP(hipMemcpy(b->p, h.data(), sizeof(float2) * (size_t)L, hipMemcpyHostToDevice));
        const float2* r = b->as<float2>();
        tw[L] = std::move(b);
        return r;
    };
    vp.x.fp.tw = table(vp.row_len());
    vp.y.fp.tw = table((int)n2);
    vp.z.fp.tw = table((int)n3);
    if (vp.packed) {
        auto r = rtw.find(n1);
        if (r == rtw.end()) {
            std::vector<float2> h;
            fill_real_twiddles(h, n1);
            auto b = std::make_unique<DevBuf>();
            alloc(*b, sizeof(float2) * h.size());
            PK_HIP(hipMemcpy(b->p, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
            r = rtw.emplace(n1, std::move(b)).first;
        }
        vp.rtw = r->second->as<float2>();
    }
    return plans.emplace(key, vp).first->second;
}

// ------------------------------------------------------------------ autotune
// Measures the candidate launch shapes of one padded size on the context's
// stream and keeps the fastest: row-pass values per thread (ratio pass), for
// each strided axis the values per thread, block size and tile width of its
// forward pass and of its convolution pass, and the pass order (which strided
// axis carries the convolution: 2 * forward(other axis) + convolution(this
// axis) per transform pair).  A measured alternative replaces the built-in
// choice only when it is at least kTuneMargin faster, so near ties keep the
// same shapes from run to run.  Buffer contents do not affect the timings: the
// workspace spectrum, I and a scratch spectrum standing in for an OTF are
// zeroed first.  Returns a one-line report.
constexpr float kTuneMargin = 0.03f;

std::string pk_ctx_impl::autotune(int64_t n1, int64_t n2, int64_t n3) {
    const std::vector<int64_t> key{n1, n2, n3};
    const VolPlan& cur = plan(n1, n2, n3);
    VolPlan vp = cur;
    const int Lx = (int)(vp.packed ? n1 / 2 : n1);
    ensure(ws.A, sizeof(float2) * (size_t)vp.S);
    ensure(ws.I, sizeof(float) * (size_t)vp.N);
    ensure(ws.tmp, sizeof(float2) * (size_t)vp.S);
    PK_HIP(hipMemsetAsync(ws.A.p, 0, sizeof(float2) * (size_t)vp.S, stream));
    PK_HIP(hipMemsetAsync(ws.I.p, 0, sizeof(float) * (size_t)vp.N, stream));
    PK_HIP(hipMemsetAsync(ws.tmp.p, 0, sizeof(float2) * (size_t)vp.S, stream));
    const bool prof_was = prof_on;
    prof_on = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PK_HIP(hipEventCreate(&e0));
    PK_HIP(hipEventCreate(&e1));
    auto timed = [&](auto&& launch) {
        launch();  // warm-up (code load, caches)
        PK_HIP(hipEventRecord(e0, stream));
        for (int r = 0; r < 3; ++r) launch();
        PK_HIP(hipEventRecord(e1, stream));
        PK_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PK_HIP(hipEventElapsedTime(&ms, e0, e1));
        return ms / 3.f;
    };
    float2* A = ws.A.as<float2>();
    const float2* H = ws.tmp.as<float2>();
    const float scale = (float)(1.0 / (double)vp.N);
    std::string rep;
    char buf[160];
    // rows
    {
        float best = 1e30f;
        AxisLaunch pick = vp.x;
        bool first = true;
        for (AxisLaunch c : x_candidates(Lx, vp.packed)) {
            c.fp.tw = vp.x.fp.tw;
            VolPlan t = vp;
            t.x = c;
            XArgs a{};
            a.mode = XM_RATIO;
            a.spec = A;
            a.I = ws.I.as<float>();
            const float ms = timed([&] { run_x(t, a); });
            std::snprintf(buf, sizeof buf, " x:E%d=%.3f", c.fp.E, ms);
            rep += buf;
            if (first || ms < best * (1.f - kTuneMargin)) {
                best = ms;
                pick = c;
            }
            first = false;
        }
        vp.x = pick;
    }
    // strided axes: best forward and best convolution shape per axis
    AxisLaunch fwd[2], conv[2];
    float tf[2], tc[2];
    for (int i = 0; i < 2; ++i) {
        const int axis = i + 2, L = (int)(axis == 2 ? n2 : n3);
        const float2* tw = (axis == 2 ? vp.y : vp.z).fp.tw;
        tf[i] = tc[i] = 1e30f;
        fwd[i] = conv[i] = axis == 2 ? vp.y : vp.z;
        std::vector<AxisLaunch> cands = col_candidates(L, axis - 1);
        if (cands.size() > (size_t)tune().col_cands) cands.resize(tune().col_cands);
        bool first = true;
        for (AxisLaunch c : cands) {
            c.fp.tw = tw;
            VolPlan t = vp;
            (axis == 2 ? t.y : t.z) = c;
            const float mf = timed([&] { run_col(t, axis, A, CM_FWD, nullptr, 1.f); });
            const float mc = timed([&] { run_col(t, axis, A, CM_CONV, H, scale); });
            std::snprintf(buf, sizeof buf, " %c:E%d/%d/T%d=%.3f/%.3f", axis == 2 ? 'y' : 'z', c.fp.E, c.nt, c.tile, mf,
                          mc);
            rep += buf;
            if (first || mf < tf[i] * (1.f - kTuneMargin)) {
                tf[i] = mf;
                fwd[i] = c;
            }
            if (first || mc < tc[i] * (1.f - kTuneMargin)) {
                tc[i] = mc;
                conv[i] = c;
            }
            first = false;
        }
    }
    // The pass order is decided on the two chains as the iteration runs them
    // (other-axis forward, convolution, other-axis inverse, back to back): the
    // sum of the kernels timed alone misjudged it (512x512x256: the dim-3
    // convolution timed 0.198 ms alone and runs 0.174 ms after the dim-2
    // forward pass).
    auto chain = [&](int cax) {
        VolPlan t = vp;
        t.y = cax == 3 ? fwd[0] : conv[0];
        t.z = cax == 3 ? conv[1] : fwd[1];
        const int oax = cax == 3 ? 2 : 3;
        return timed([&] {
            run_col(t, oax, A, CM_FWD, nullptr, 1.f);
            run_col(t, cax, A, CM_CONV, H, scale);
            run_col(t, oax, A, CM_INV, nullptr, 1.f);
        });
    };
    const float order_z = chain(3), order_y = chain(2);
    if (order_y < order_z * (1.f - kTuneMargin)) {  // the built-in order convolves along dim 3
        vp.conv_axis = 2;
        vp.y = conv[0];
        vp.z = fwd[1];
    } else {
        vp.conv_axis = 3;
        vp.y = fwd[0];
        vp.z = conv[1];
    }
    std::snprintf(buf, sizeof buf, " | conv along dim %d (%.3f vs %.3f ms per transform pair)", vp.conv_axis,
                  std::min(order_y, order_z), std::max(order_y, order_z));
    rep += buf;
    vp.tuned = true;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    prof_on = prof_was;
    plans[key] = vp;
    clear_graphs();  // captured loops hold the old launch shapes
    {
        Wisdom& w = Wisdom::get();
        std::lock_guard<std::mutex> lk(w.mu);
        w.store(key, {vp.x.fp.E, vp.y.fp.E, vp.y.nt, vp.y.tile, vp.z.fp.E, vp.z.nt, vp.z.tile, vp.conv_axis});
    }
    std::snprintf(buf, sizeof buf, "%lldx%lldx%lld: x E%d; y E%d/%d/T%d; z E%d/%d/T%d; conv dim %d |",
                  (long long)n1, (long long)n2, (long long)n3, vp.x.fp.E, vp.y.fp.E, vp.y.nt, vp.y.tile, vp.z.fp.E,
                  vp.z.nt, vp.z.tile, vp.conv_axis);
    return std::string(buf) + rep;
}

int pk_ctx_impl::chunk_streams() const { return tune().streams; }

// ------------------------------------------------------------------ graphs
void pk_ctx_impl::clear_graphs() {
    for (auto& g : graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

template <class Fn>
void pk_ctx_impl::run_graphed(const std::vector<uint64_t>& key, Fn&& fn) {
    // not while profiling (per-pass events), not on the legacy null stream, and PK_GRAPHS=0 disables
    if (!tune().graphs || prof_on || !stream) {
        fn();
        return;
    }
    GraphEntry* e = nullptr;
    for (auto& g : graphs)
        if (g.key == key) e = &g;
    if (!e) {  // first sighting: eager (code objects loaded, attributes set, plan verified)
        if (graphs.size() >= 16) {
            auto lru = std::min_element(graphs.begin(), graphs.end(),
                                        [](const GraphEntry& a, const GraphEntry& b) { return a.tick < b.tick; });
            if (lru->exec) (void)hipGraphExecDestroy(lru->exec);
            graphs.erase(lru);
        }
        graphs.push_back(GraphEntry{key, nullptr, ++graph_tick});
        fn();
        return;
    }
    e->tick = ++graph_tick;
    if (!e->exec) {
        PK_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            fn();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        PK_HIP(hipStreamEndCapture(stream, &g));
        const hipError_t ie = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PK_HIP(ie);
    }
    PK_HIP(hipGraphLaunch(e->exec, stream));
}

// ------------------------------------------------------------------ passes
void pk_ctx_impl::run_x(const VolPlan& vp, XArgs a, int64_t z0, int64_t nz) {
    a.n1 = (int)vp.n1;
    a.h1 = (int)vp.h1;
    a.hp = (int)vp.hp;
    a.packed = vp.packed;
    a.RT = vp.x.tile;
    a.lpad = vp.x.lpad;
    a.rtw = vp.rtw;
    a.eps = kEps;
    if (nz <= 0) nz = vp.n3 - z0;
    a.nrows = vp.n2 * nz;
    if (z0 > 0) {
        // slab: every array advanced to plane z0; block partials continue the
        // whole-volume numbering (slab boundaries fall on block boundaries)
        const int64_t ro = z0 * vp.n2 * vp.n1;
        auto adv = [&](auto*& ptr) {
            if (ptr) ptr += ro;
        };
        adv(a.src);
        adv(a.dst);
        adv(a.raw16);
        adv(a.rawf);
        adv(a.I);
        adv(a.I16);
        adv(a.J);
        adv(a.Jin);
        adv(a.J3);
        adv(a.Y);
        adv(a.Jn);
        adv(a.J4);
        a.spec += z0 * vp.n2 * vp.hp;
        if (a.partials) a.partials += 2 * (z0 * vp.n2 / a.RT);
    }
    const double frac = (double)nz / (double)vp.n3;
    const int64_t nblk = (a.nrows + a.RT - 1) / a.RT;
    hipEvent_t e0 = prof_mark();
    // compile-time kernel when one exists for this (row length, values per
    // thread) and the tile is the full 256-thread block it assumes
    if (vp.x.block.x == (unsigned)kMaxThreads && !tune().generic_x &&
        launch_xrow(vp.packed ? vp.x.fp.L : -vp.x.fp.L, vp.x.fp.E, dim3((unsigned)nblk), vp.x.lds, stream, a,
                    vp.x.fp.tw))
        PK_HIP(hipGetLastError());
    else if (xm_u16(a.mode))
        fail(PK_ERR_INVALID_ARG, "uint16-operand row modes need the compile-time row kernel");
    else
        PK_HIP(launch_xpass(vp.x.fp, dim3((unsigned)nblk), vp.x.block, vp.x.lds, stream, a));
    if (prof_on) {
        // algorithmic bytes: each operand read or written once
        const double N = (double)vp.N, S = (double)vp.Sa;
        const char* name = "x_pass";
        double b = 0;
        switch (a.mode) {
            case XM_R2C: name = "x_r2c"; b = 4 * N + 8 * S; break;
            case XM_C2R: name = "x_c2r"; b = 8 * S + 4 * N; break;
            case XM_PREP_U16: name = "x_prep"; b = 2 * N + (a.I16 ? 2 : 4) * N + (a.J ? 4 : 0) * N + 8 * S; break;
            case XM_PREP_F32: name = "x_prep"; b = 4 * N + (a.J ? 8 : 4) * N + 8 * S; break;
            case XM_RATIO: name = "x_ratio"; b = 16 * S + 4 * N; break;
            case XM_RATIO_U16: name = "x_ratio"; b = 16 * S + 2 * N; break;
            case XM_OMW_UPDATE_U16: name = "x_update_r2c"; b = 8 * S + 6 * N + 8 * S; break;
            case XM_OMW_UPDATE: name = a.r2c_out ? "x_update_r2c" : "x_update"; b = 8 * S + 8 * N + (a.r2c_out ? 8 * S : 0); break;
            case XM_BIGGS_Y: name = "x_biggs_y"; b = 12 * N + 8 * S; break;
            case XM_BIGGS_UPDATE: name = "x_biggs_update"; b = 8 * S + 12 * N + (a.acc ? 4 * N : 0) + (a.r2c_out ? 8 * S : 0); break;
        }
        if (a.mode == XM_OMW_FINAL || a.mode == XM_BIGGS_FINAL) {
            // fused last update: spectrum and estimate operand in, the
            // epilogue's output (and damp / mask operands) out; no J store
            const EpiArgs& ep = a.ep;
            const double on = (double)ep.o1 * ep.o2 * ep.o3;
            const double mb = (ep.apply_scale && ep.use_mask) ? (ep.mask ? 1.0 : ep.raw16 ? 2.0 : 4.0) : 0.0;
            name = a.mode == XM_OMW_FINAL ? "x_update_epi" : "x_biggs_update_epi";
            b = 8 * S + 4 * N + on * ((ep.out16 ? 2.0 : 4.0) + (ep.damp > 1.f ? 4.0 : 0.0) + mb);
        }
        prof_done(name, b * frac, e0);
    }
}

static int64_t xpass_blocks(const VolPlan& vp) { return (vp.n2 * vp.n3 + vp.x.tile - 1) / vp.x.tile; }

void pk_ctx_impl::run_col(const VolPlan& vp, int axis, float2* data, int mode, const float2* otf, float scale,
                          int64_t z0, int64_t nz) {
    const AxisLaunch& ax = axis == 2 ? vp.y : vp.z;
    if (nz <= 0) nz = vp.n3 - z0;
    if (axis == 2 && z0 > 0) data += z0 * vp.hp * vp.n2;  // dim-2 columns of planes [z0, z0 + nz)
    if (ax.fp.L == 1 && (mode == CM_FWD || mode == CM_INV)) return;  // length-1 transform: identity
    ColArgs a{};
    a.data = data;
    a.otf = otf;
    a.T = ax.tile;
    a.mode = mode;
    a.scale = scale;
    dim3 grid;
    if (axis == 2) {
        a.acount = vp.h1;
        a.bstride = vp.hp * vp.n2;
        a.stride = vp.hp;
        grid = dim3((unsigned)((vp.h1 + a.T - 1) / a.T), (unsigned)nz);
    } else {
        a.acount = vp.hp * vp.n2;
        a.amod = vp.hp;
        a.alim = vp.h1;
        a.bstride = 0;
        a.stride = vp.hp * vp.n2;
        grid = dim3((unsigned)((a.acount + a.T - 1) / a.T), 1);
    }
    // the remap needs a whole number of tiles per XCD
    a.xcd = (tune().xcd && ((int64_t)grid.x * grid.y) % 8 == 0) ? 1 : 0;
    hipEvent_t e0 = prof_mark();
    PK_HIP(launch_colpass(ax.fp, ax.nt, grid, ax.block, ax.lds, stream, a));
    if (prof_on) {
        const double S = (double)vp.Sa;
        const char* name = axis == 2 ? (mode == CM_INV ? "y_inv" : mode == CM_FWD ? "y_fwd" : "y_conv")
                                     : (mode == CM_INV ? "z_inv" : mode == CM_FWD ? "z_fwd" : "z_conv");
        const double frac = axis == 2 ? (double)nz / (double)vp.n3 : 1.0;
        prof_done(name, (mode >= CM_CONV ? 24 * S : 16 * S) * frac, e0);
    }
}

void pk_ctx_impl::r2c3d(const VolPlan& vp, const float* real, float2* spec) {
    XArgs a{};
    a.mode = XM_R2C;
        a.src = real;
    a.spec = spec;
    run_x(vp, a);
    run_col(vp, 2, spec, CM_FWD, nullptr, 1.f);
    run_col(vp, 3, spec, CM_FWD, nullptr, 1.f);
}

void pk_ctx_impl::c2r3d(const VolPlan& vp, float2* spec, float* real) {
    run_col(vp, 3, spec, CM_INV, nullptr, 1.f);
    run_col(vp, 2, spec, CM_INV, nullptr, 1.f);
    XArgs a{};
    a.mode = XM_C2R;
        a.spec = spec;
    a.dst = real;
    run_x(vp, a);
}

static int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 65536); }

// Planes per slab for the dim-1 / dim-2 passes: the chain (dim-2 inverse,
// dim-1 step, dim-2 forward) runs slab by slab so its working set stays in the
// device's memory-side cache; 0 = whole volume.  A slab holds whole x-pass
// blocks (nz * n2 a multiple of the rows per block).  PK_SLAB_PLANES overrides.
bool pk_ctx_impl::x_compile_time(const VolPlan& vp) const {
    return vp.x.block.x == (unsigned)kMaxThreads && !tune().generic_x &&
           xrow_available(vp.packed ? vp.x.fp.L : -vp.x.fp.L, vp.x.fp.E);
}

int64_t pk_ctx_impl::slab_planes(const VolPlan& vp) const {
    int64_t k = tune().slab >= 0 ? tune().slab : 0;
    if (k <= 0 || k >= vp.n3 || vp.conv_axis != 3) return 0;
    while (k < vp.n3 && (k * vp.n2) % vp.x.tile != 0) ++k;
    return k >= vp.n3 ? 0 : k;
}

// ------------------------------------------------------------------ OTF cache
// 64-bit content hash of a PSF (cache key): four independent multiply-xorshift
// lanes over 8-byte words, ~1 ns per 32 bytes, then a final avalanche.
static inline uint64_t mix64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}
static uint64_t content_hash(const void* p, size_t n, uint64_t seed) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    uint64_t h[4] = {seed ^ 0x9e3779b97f4a7c15ull, seed + 0x632be59bd9b4e019ull, seed ^ 0x85ebca77c2b2ae63ull,
                     seed + 0x27d4eb2f165667c5ull};
    size_t i = 0;
    for (; i + 32 <= n; i += 32) {
        uint64_t w[4];
        std::memcpy(w, c + i, 32);
        for (int l = 0; l < 4; ++l) {
            h[l] = (h[l] ^ w[l]) * 0x9e3779b97f4a7c15ull;
            h[l] ^= h[l] >> 29;
        }
    }
    uint64_t t = 0;
    for (; i < n; ++i) t = (t << 8) ^ c[i] ^ (t >> 56);
    return mix64(h[0] ^ mix64(h[1] ^ mix64(h[2] ^ mix64(h[3] ^ t ^ n))));
}

// OTF of PSF ./ sum(PSF(:)) at the volume size (decon_lucy_function.m:76,
// decon_lucy_omw_function.m:70-71), half spectrum, cached on the device.
uint64_t psf_key(const float* psf, const int64_t pd[3]) {
    const int64_t P = pd[0] * pd[1] * pd[2];
    return content_hash(psf, sizeof(float) * (size_t)P, content_hash(pd, sizeof(int64_t) * 3, 0x5eed));
}

uint64_t key_mix(uint64_t h, const int64_t* v, int n) { return content_hash(v, sizeof(int64_t) * (size_t)n, h); }

const float2* pk_ctx_impl::otf_for(const VolPlan& vp, const float* psf, const int64_t pd[3]) {
    const uint64_t h = psf_key(psf, pd);
    if (const float2* hit = otf_find(vp, h, pd)) return hit;
    return otf_build(vp, h, psf, pd);
}

const float2* pk_ctx_impl::otf_find(const VolPlan& vp, uint64_t h, const int64_t pd[3]) {
    for (auto& e : otf) {
        if (e.dims[0] == vp.n1 && e.dims[1] == vp.n2 && e.dims[2] == vp.n3 && e.hash == h &&
            e.pdims[0] == pd[0] && e.pdims[1] == pd[1] && e.pdims[2] == pd[2]) {
            e.tick = ++tick;
            return e.spec->as<float2>();
        }
    }
    return nullptr;
}

const float2* pk_ctx_impl::otf_build(const VolPlan& vp, uint64_t h, const float* psf, const int64_t pd[3]) {
    const int64_t P = pd[0] * pd[1] * pd[2];
    if (pd[0] > vp.n1 || pd[1] > vp.n2 || pd[2] > vp.n3)
        fail(PK_ERR_OUT_SIZE, "outSizeIsSmallerThanPsfSize: PSF larger than the volume");
    // evict least recently used among the unpinned entries
    while (otf.size() >= otf_max_entries) {
        auto lru = otf.end();
        for (auto e = otf.begin(); e != otf.end(); ++e)
            if (e->pins == 0 && (lru == otf.end() || e->tick < lru->tick)) lru = e;
        if (lru == otf.end()) break;  // all in use: grow
        otf.erase(lru);
    }
    // PSF ./ sum(PSF(:)): sum accumulated in double, divided in single
    std::vector<float> q((size_t)P);
    double s = 0;
    for (int64_t i = 0; i < P; ++i) s += (double)psf[i];
    const float sf = (float)s;
    for (int64_t i = 0; i < P; ++i) q[(size_t)i] = psf[i] / sf;

    OtfEntry ent{};
    ent.dims[0] = vp.n1;
    ent.dims[1] = vp.n2;
    ent.dims[2] = vp.n3;
    std::memcpy(ent.pdims, pd, sizeof(int64_t) * 3);
    ent.hash = h;
    ent.tick = ++tick;
    ent.spec = std::make_unique<DevBuf>();
    alloc(*ent.spec, sizeof(float2) * (size_t)vp.S);
    build_otf(vp, q.data(), pd, ent.spec->as<float2>());
    otf.push_back(std::move(ent));
    return otf.back().spec->as<float2>();
}

void pk_ctx_impl::otf_pin(const float2* spec, int delta) {
    for (auto& e : otf)
        if (e.spec->as<float2>() == spec) {
            e.pins += delta;
            return;
        }
}

void pk_ctx_impl::build_otf(const VolPlan& vp, const float* q, const int64_t pd[3], float2* dst) {
    const int64_t P = pd[0] * pd[1] * pd[2];
    ensure(ws.tmp, sizeof(float) * (size_t)std::max<int64_t>(vp.N, P));
    DevBuf dq;
    alloc(dq, sizeof(float) * (size_t)P);
    build_otf_into(vp, q, pd, dst, ws.tmp.as<float>(), dq.as<float>());
    PK_HIP(hipStreamSynchronize(stream));  // dq is freed on return
    prof_flush();
}

void pk_ctx_impl::build_otf_into(const VolPlan& vp, const float* q, const int64_t pd[3], float2* dst, float* real,
                                 float* dq) {
    const int64_t P = pd[0] * pd[1] * pd[2];
    PK_HIP(hipMemcpyAsync(dq, q, sizeof(float) * (size_t)P, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(vp.N)), dim3(256), 0, stream, real, vp.N, 0.f);
    PK_HIP(hipGetLastError());
    bool allzero = true;
    for (int64_t i = 0; i < P && allzero; ++i) allzero = (q[i] == 0.f);
    if (!allzero) {  // an all-zero PSF gives a zero OTF (decon_psf2otf.m:42-44)
        hipLaunchKernelGGL(k_scatter_psf, dim3(grid_for(P)), dim3(256), 0, stream, dq, (int)pd[0], (int)pd[1],
                           (int)pd[2], real, vp.n1, vp.n2, vp.n3);
        PK_HIP(hipGetLastError());
    }
    r2c3d(vp, real, dst);
}

void pk_ctx_impl::reset() {
    PK_HIP(hipStreamSynchronize(stream));
    prof_flush();
    clear_graphs();
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    ev_pool.clear();
    otf.clear();
    plans.clear();
    tw.clear();
    rtw.clear();
    ws.I.release(); ws.J.release(); ws.A.release(); ws.J3.release(); ws.Y.release(); ws.J4.release();
    ws.partials.release(); ws.lambda.release(); ws.raw.release(); ws.out.release(); ws.mask.release();
    ws.mask2.release(); ws.count.release(); ws.tmp.release(); ws.batch.release(); ws.dsr.release(); ws.dbg.release();
    block_cache_clear();
    for (int i = 0; i < 2; ++i) {
        ws.tab_dev[i].release();
        if (ws.tab_host[i].p) (void)hipHostFree(ws.tab_host[i].p);
        ws.tab_host[i].p = nullptr;
        ws.tab_host[i].bytes = 0;
        if (ws.tab_ev[i]) (void)hipEventDestroy(ws.tab_ev[i]);
        ws.tab_ev[i] = nullptr;
    }
}

// ------------------------------------------------------------------ decon driver
static double matlab_u16(double x) {
    if (!(x > 0)) return 0;
    if (x >= 65535.0) return 65535;
    return std::round(x);
}

DeconGeom decon_check(int method, const pk_vol* Iv, const float* psf_f, const int64_t fd[3], const float* psf_b,
                      const int64_t bd[3], int numit, const pk_opts* o, const pk_vol* out) {
    if (!Iv || !out || !o || !psf_f || numit < 0) fail(PK_ERR_INVALID_ARG, "null argument or negative NUMIT");
    if (method == 1 && !psf_b) fail(PK_ERR_INVALID_ARG, "PSF_b missing");
    if (Iv->dtype != PK_UINT16 && Iv->dtype != PK_SINGLE)
        fail(PK_ERR_UNSUPPORTED_TYPE, "input class must be uint16 or single");
    const int64_t n1 = Iv->dims[0], n2 = Iv->dims[1], n3 = Iv->dims[2];
    if (n1 < 1 || n2 < 1 || n3 < 1) fail(PK_ERR_INVALID_ARG, "empty volume");
    for (int d = 0; d < 3; ++d) {
        if (fd[d] < 1 || (method == 1 && bd[d] < 1)) fail(PK_ERR_INVALID_ARG, "empty PSF");
        if (fd[d] > Iv->dims[d] || (method == 1 && bd[d] > Iv->dims[d]))
            fail(PK_ERR_OUT_SIZE, "outSizeIsSmallerThanPsfSize: PSF larger than the volume");
    }
    // output geometry (bbox, 1-based inclusive)
    DeconGeom g{{0, 0, 0}, {n1, n2, n3}};
    if (o->has_bbox) {
        for (int d = 0; d < 3; ++d) {
            const int64_t s = o->bbox[d], e = o->bbox[d + 3];
            if (s < 1 || e > Iv->dims[d] || e < s) fail(PK_ERR_INVALID_ARG, "bbox outside the volume");
            g.b0[d] = s - 1;
            g.od[d] = e - s + 1;
        }
    }
    if (out->dims[0] != g.od[0] || out->dims[1] != g.od[1] || out->dims[2] != g.od[2])
        fail(PK_ERR_INVALID_ARG, "output dims do not match the (bbox) result size");
    if (out->dtype != (o->save16bit ? PK_UINT16 : PK_SINGLE))
        fail(PK_ERR_INVALID_ARG, "output class must be uint16 when save16bit, single otherwise");
    if (!out->data || !Iv->data) fail(PK_ERR_INVALID_ARG, "null data pointer");
    return g;
}

// decon_mask_edge_erosion(I ~= 0, E) (decon_mask_edge_erosion.m:1-30) on the
// stream: m1, m2 (N bytes each) and cnt (two counters) are the caller's;
// returns the buffer holding the eroded mask.
const uint8_t* pk_ctx_impl::edge_mask(const uint16_t* raw16, const float* rawf, int64_t n1, int64_t n2, int64_t n3,
                                      int E, uint8_t* m1, uint8_t* m2, unsigned long long* cnt) {
    const int64_t N = n1 * n2 * n3;
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, stream, reinterpret_cast<float*>(cnt), (int64_t)4, 0.f);
    const int is3d = n3 > 1;
    hipLaunchKernelGGL(k_mask_init, dim3(grid_for(N)), dim3(256), 0, stream, raw16, rawf, N, m1, cnt);
    hipLaunchKernelGGL(k_mask_faces, dim3(grid_for(N)), dim3(256), 0, stream, m1, n1, n2, n3, is3d, cnt, E);
    hipLaunchKernelGGL(k_mask_erode_disk, dim3(grid_for(N)), dim3(256), 0, stream, m1, m2, n1, n2, n3, disk_se(E - 1),
                       cnt);
    const uint8_t* mask = m2;
    if (is3d) {
        hipLaunchKernelGGL(k_mask_erode_z, dim3(grid_for(N)), dim3(256), 0, stream, m2, m1, n1, n2, n3, E - 1, cnt);
        mask = m1;
    }
    PK_HIP(hipGetLastError());
    return mask;
}

EpiArgs epi_args(const VolPlan& vp, const pk_opts* o, const DeconGeom& g, const float* I, const uint8_t* mask,
                 const uint16_t* raw16, const float* rawf, void* dout) {
    const size_t osz = o->save16bit ? 2 : 4;
    EpiArgs e{};
    e.I = I;
    e.mask = mask;
    e.raw16 = raw16;
    e.rawf = rawf;
    e.out16 = o->save16bit ? static_cast<uint16_t*>(dout) : nullptr;
    e.outf = o->save16bit ? nullptr : static_cast<float*>(dout);
    e.n1 = vp.n1;
    e.n2 = vp.n2;
    e.o1 = g.od[0];
    e.o2 = g.od[1];
    e.o3 = g.od[2];
    e.b1 = g.b0[0];
    e.b2 = g.b0[1];
    e.b3 = g.b0[2];
    e.damp = (float)o->damp_factor;
    e.scale = (float)o->scale_factor;
    e.offset = (float)o->decon_offset;
    e.apply_scale = (o->scale_factor != 1 || o->decon_offset != 0 || o->edge_erosion > 0);
    e.use_mask = o->decon_offset != 0 || o->edge_erosion > 0;
    if (!(o->damp_factor > 1)) e.damp = 0.f;
    // pair stores in the fused form: even dim-1 crop start and size, aligned output
    e.pair = vp.packed && g.b0[0] % 2 == 0 && g.od[0] % 2 == 0 &&
             reinterpret_cast<uintptr_t>(dout) % (2 * osz) == 0;
    return e;
}

// prep (decon_lucy_function.m:85-100, GPU branch): uint16 I - uint16(Background)
// saturates at 0; single I gives max(I - Background, 0).  Writes I, its
// half spectrum into A and (Biggs) J_2 = I.
static XArgs prep_args(int method, float2* A, float* I, float* J, const uint16_t* raw16, const float* rawf,
                       double background) {
    XArgs pa{};
    pa.spec = A;
    pa.I = I;
    pa.J = method == 1 ? nullptr : J;  // OMW: the first update reads I (J_1 = I) instead of a copy
    if (raw16) {
        pa.mode = XM_PREP_U16;
        pa.raw16 = raw16;
        pa.bg = (float)matlab_u16(background);
    } else {
        pa.mode = XM_PREP_F32;
        pa.rawf = rawf;
        pa.bg = (float)background;
    }
    return pa;
}

// OMW passes (decon_lucy_omw_function.m:105-112) after nothing but the OTFs:
// prep, then per iteration J_{k+1} = max(J_k .* real(ifftn(fftn(I ./
// max(real(ifftn(fftn(J_k) .* OTF_f)), eps)) .* OTF_b)), 0); with fuse_epi
// the last update writes the output through the epilogue e.  The dim-1 and
// dim-2 passes run slab by slab (kz planes; 0 = the whole volume).  loop_key:
// the iteration loop is replayed as a graph for a repeated key (null: eager).
void pk_ctx_impl::omw_passes(const VolPlan& vp, const XArgs& pa, float* I, float* J, float2* A, const float2* Hf,
                             const float2* Hb, int numit, const EpiArgs& e, bool fuse_epi, int64_t kz,
                             const std::vector<uint64_t>* loop_key,
                             const std::function<void(int, const float*)>* after, uint16_t* I16) {
    const int64_t step = kz > 0 ? kz : vp.n3;
    const float scale = (float)(1.0 / (double)vp.N);  // ifftn's 1/N
    auto slabs = [&](auto&& body) {
        for (int64_t z0 = 0; z0 < vp.n3; z0 += step) body(z0, std::min(step, vp.n3 - z0));
    };
    // the strided pass pair around the convolution (dim 2 unless the plan
    // convolves along dim 2); slabs only split dim-2 passes
    const int outer = vp.conv_axis == 3 ? 2 : 3;
    auto y_fwd = [&](int64_t z0, int64_t nz) { run_col(vp, outer, A, CM_FWD, nullptr, 1.f, z0, nz); };
    auto y_inv = [&](int64_t z0, int64_t nz) { run_col(vp, outer, A, CM_INV, nullptr, 1.f, z0, nz); };
    auto z_conv = [&](const float2* H) { run_col(vp, vp.conv_axis, A, CM_CONV, H, scale); };
    slabs([&](int64_t z0, int64_t nz) {
        run_x(vp, pa, z0, nz);
        if (numit > 0) y_fwd(z0, nz);
    });
    auto loop = [&] {
        for (int k = 1; k <= numit; ++k) {
            z_conv(Hf);
            slabs([&](int64_t z0, int64_t nz) {
                y_inv(z0, nz);
                XArgs r{};
                r.mode = I16 ? XM_RATIO_U16 : XM_RATIO;
                r.spec = A;
                r.I = I;
                r.I16 = I16;
                run_x(vp, r, z0, nz);
                y_fwd(z0, nz);
            });
            z_conv(Hb);
            slabs([&](int64_t z0, int64_t nz) {
                y_inv(z0, nz);
                XArgs u{};
                u.mode = (k == 1 && I16) ? XM_OMW_UPDATE_U16 : XM_OMW_UPDATE;  // J_1 = I, from I16
                u.spec = A;
                u.J = J;
                u.Jin = k == 1 ? I : nullptr;
                u.I16 = I16;
                u.r2c_out = k < numit;
                if (!(fuse_epi && k == numit)) run_x(vp, u, z0, nz);
                if (k < numit) y_fwd(z0, nz);
            });
            if (after) (*after)(k, J);
        }
    };
    if (loop_key && !after) run_graphed(*loop_key, loop);
    else loop();
    if (fuse_epi) {  // last update + epilogue
        XArgs u{};
        u.mode = XM_OMW_FINAL;
        u.spec = A;
        u.J = J;
        u.Jin = numit == 1 ? I : nullptr;
        u.ep = e;
        run_x(vp, u);
    }
}

void pk_ctx_impl::decon(int method, const pk_vol* Iv, const float* psf_f, const int64_t fd[3], const float* psf_b,
                        const int64_t bd[3], int numit, const pk_opts* o, pk_vol* out, double* err_mat) {
    const DeconGeom g = decon_check(method, Iv, psf_f, fd, psf_b, bd, numit, o, out);
    const int64_t n1 = Iv->dims[0], n2 = Iv->dims[1], n3 = Iv->dims[2];
    const int64_t* od = g.od;

    if ((o->flags & PK_FLAG_AUTOTUNE) && !plan(n1, n2, n3).tuned) autotune(n1, n2, n3);
    const VolPlan& vp = plan(n1, n2, n3);
    const int64_t N = vp.N;
    if (o->flags & PK_FLAG_NO_OTF_CACHE) {  // rebuild the OTFs on every call (comparison baseline)
        otf.erase(std::remove_if(otf.begin(), otf.end(), [](const OtfEntry& e) { return e.pins == 0; }), otf.end());
    }
    const float2* Hf = otf_for(vp, psf_f, fd);
    const float2* Hb = method == 1 ? otf_for(vp, psf_b, bd) : nullptr;

    // workspace
    ensure(ws.I, sizeof(float) * N);
    ensure(ws.J, sizeof(float) * N);
    ensure(ws.A, sizeof(float2) * vp.S);
    const int64_t nxb = xpass_blocks(vp);
    if (method == 0) {
        ensure(ws.J3, sizeof(float) * N);
        ensure(ws.Y, sizeof(float) * N);
        ensure(ws.J4, sizeof(float) * N);
        ensure(ws.partials, sizeof(double) * 2 * (size_t)std::max<int64_t>(nxb, 1024));  // 1024: unfused k_dot2
        ensure(ws.lambda, sizeof(float) * 4);
    }
    const size_t esz = Iv->dtype == PK_UINT16 ? 2 : 4;
    const void* raw = Iv->data;
    if (Iv->mem == PK_MEM_HOST) {
        ensure(ws.raw, esz * N);
        PK_HIP(hipMemcpyAsync(ws.raw.p, Iv->data, esz * N, hipMemcpyHostToDevice, stream));
        raw = ws.raw.p;
    }
    const uint16_t* raw16 = Iv->dtype == PK_UINT16 ? static_cast<const uint16_t*>(raw) : nullptr;
    const float* rawf = Iv->dtype == PK_SINGLE ? static_cast<const float*>(raw) : nullptr;

    float* I = ws.I.as<float>();
    float* J = ws.J.as<float>();
    float2* A = ws.A.as<float2>();
    // uint16 input: I = sat(raw - bg16) holds integers below 2^16, so the prep
    // writes it as uint16 (2 bytes per voxel instead of 4) and the ratio and the
    // first OMW update read it so, with the same bits.  Float I where something
    // else reads it: the damp epilogue, the debug metric, OMW with one iteration
    // (its fused last update), no iteration (J = I) and the unfused path; and
    // only where the row passes have compile-time kernels (the uint16 modes).
    // Float I is then neither written nor read, so the uint16 copy takes the
    // first half of its buffer.
    uint16_t* I16 = nullptr;
    if (raw16 && numit >= (method == 1 ? 2 : 1) && !(o->flags & (PK_FLAG_UNFUSED | PK_FLAG_DEBUG)) &&
        !(o->damp_factor > 1) && x_compile_time(vp))
        I16 = static_cast<uint16_t*>(ws.I.p);

    // epilogue (decon_lucy_function.m:177-203)
    const uint8_t* mask = nullptr;
    if (o->edge_erosion > 0) {
        ensure(ws.mask, (size_t)N);
        ensure(ws.mask2, (size_t)N);
        ensure(ws.count, 32);
        mask = edge_mask(raw16, rawf, n1, n2, n3, (int)o->edge_erosion, ws.mask.as<uint8_t>(), ws.mask2.as<uint8_t>(),
                         ws.count.as<unsigned long long>());
    }
    const int64_t on = od[0] * od[1] * od[2];
    const size_t osz = o->save16bit ? 2 : 4;
    void* dout = out->data;
    if (out->mem == PK_MEM_HOST) {
        ensure(ws.out, osz * on);
        dout = ws.out.p;
    }
    EpiArgs e = epi_args(vp, o, g, I, mask, raw16, rawf, dout);
    // the last update writes the output itself (fused epilogue) when the row
    // pass covers the whole volume in one launch
    const bool dbg = (o->flags & PK_FLAG_DEBUG) != 0;
    if (dbg && (o->flags & PK_FLAG_UNFUSED)) fail(PK_ERR_INVALID_ARG, "PK_FLAG_DEBUG runs the fused engine only");
    const int64_t kz = (o->flags & PK_FLAG_UNFUSED) ? 0 : slab_planes(vp);
    const bool fuse_epi = !(o->flags & (PK_FLAG_UNFUSED | PK_FLAG_EPILOGUE_PASS)) && numit > 0 && kz == 0 && !dbg;
    // debug: MSE of every iteration's estimate against I (double sums), the hook every dbg_step iterations
    std::vector<double> mse;
    std::vector<float> snap;
    const std::function<void(int, const float*)> dbg_after = [&](int k, const float* Jk) {
        const int nb = 1024;
        ensure(ws.dbg, sizeof(double) * (size_t)(nb + 1));
        double* part = ws.dbg.as<double>();
        hipLaunchKernelGGL(k_sqdiff_partial, dim3(nb), dim3(256), 0, stream, Jk, (const float*)I, N, part);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, stream, (const double*)part, nb, part + nb);
        PK_HIP(hipGetLastError());
        double sum = 0;
        PK_HIP(hipMemcpyAsync(&sum, part + nb, sizeof(double), hipMemcpyDeviceToHost, stream));
        PK_HIP(hipStreamSynchronize(stream));
        mse.push_back(sum / (double)N);
        if (dbg_hook && dbg_step > 0 && k % dbg_step == 0) {
            snap.resize((size_t)N);
            PK_HIP(hipMemcpy(snap.data(), Jk, sizeof(float) * (size_t)N, hipMemcpyDeviceToHost));
            const int64_t dd[3] = {vp.n1, vp.n2, vp.n3};
            dbg_hook(dbg_user, k, snap.data(), dd);
        }
    };
    if (o->flags & PK_FLAG_UNFUSED) {
        const float bgv = raw16 ? (float)matlab_u16(o->background) : (float)o->background;
        decon_unfused(method, vp, raw, raw16 != nullptr, bgv, Hf, Hb, numit, I, J);
    } else {
        // The dim-1 and dim-2 passes run slab by slab (planes [z0, z0 + kz)):
        // each slab's chain (dim-2 inverse, dim-1 step, dim-2 forward) reuses
        // its data from the memory-side cache; the dim-3 convolution covers the
        // whole volume between chains.  kz = 0: one slab (the whole volume).
        const int64_t step = kz > 0 ? kz : vp.n3;
        const float scale = (float)(1.0 / (double)vp.N);  // ifftn's 1/N
        auto slabs = [&](auto&& body) {
            for (int64_t z0 = 0; z0 < vp.n3; z0 += step) body(z0, std::min(step, vp.n3 - z0));
        };
        const int outer = vp.conv_axis == 3 ? 2 : 3;
        auto y_fwd = [&](int64_t z0, int64_t nz) { run_col(vp, outer, A, CM_FWD, nullptr, 1.f, z0, nz); };
        auto y_inv = [&](int64_t z0, int64_t nz) { run_col(vp, outer, A, CM_INV, nullptr, 1.f, z0, nz); };
        auto z_conv = [&](const float2* H, bool conj) {
            run_col(vp, vp.conv_axis, A, conj ? CM_CONVJ : CM_CONV, H, scale);
        };

        // the iteration loop reads only the workspace, the cached OTFs and the
        // plan's tables, so it is replayed as a graph for a repeated key
        auto u64 = [](const void* p) { return (uint64_t)reinterpret_cast<uintptr_t>(p); };
        const std::vector<uint64_t> loop_key{
            (uint64_t)method, (uint64_t)numit, u64(&vp), (uint64_t)vp.N, u64(Hf), u64(Hb), u64(ws.I.p), u64(ws.J.p),
            u64(ws.A.p), u64(ws.J3.p), u64(ws.Y.p), u64(ws.J4.p), u64(ws.lambda.p), u64(ws.partials.p),
            (uint64_t)kz, u64(stream), (uint64_t)fuse_epi, u64(I16)};
        XArgs pa = prep_args(method, A, I, J, raw16, rawf, o->background);
        pa.I16 = I16;

        if (method == 1) {
            omw_passes(vp, pa, I, J, A, Hf, Hb, numit, e, fuse_epi, kz, &loop_key, dbg ? &dbg_after : nullptr, I16);
        } else {
            run_x(vp, pa);
            float* J2 = J;
            float* J3 = ws.J3.as<float>();
            float* Y = ws.Y.as<float>();
            float* J4 = ws.J4.as<float>();
            float* lam = ws.lambda.as<float>();
            double* part = ws.partials.as<double>();
            auto biggs_loop = [&] {
            J2 = J;
            J3 = ws.J3.as<float>();
            hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(N)), dim3(256), 0, stream, J3, N, 0.f);
            hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, stream, lam, (int64_t)1, 0.f);
            PK_HIP(hipGetLastError());
            for (int k = 1; k <= numit; ++k) {
                if (k > 2) {  // lambda from the dot products of every slab of the last update
                    hipLaunchKernelGGL(k_lambda, dim3(1), dim3(256), 0, stream, part, (int)nxb, kEps, lam);
                    PK_HIP(hipGetLastError());
                }
                slabs([&](int64_t z0, int64_t nz) {
                    XArgs y{};
                    y.mode = XM_BIGGS_Y;
                    y.spec = A;
                    y.J = J2;
                    y.J3 = J3;
                    y.Y = Y;
                    y.lambda = lam;
                    run_x(vp, y, z0, nz);
                    y_fwd(z0, nz);
                });
                z_conv(Hf, false);
                slabs([&](int64_t z0, int64_t nz) {
                    y_inv(z0, nz);
                    XArgs r{};
                    r.mode = I16 ? XM_RATIO_U16 : XM_RATIO;
                    r.spec = A;
                    r.I = I;
                    r.I16 = I16;
                    run_x(vp, r, z0, nz);
                    y_fwd(z0, nz);
                });
                z_conv(Hf, true);
                slabs([&](int64_t z0, int64_t nz) {
                    y_inv(z0, nz);
                    XArgs u{};
                    u.mode = XM_BIGGS_UPDATE;
                    u.spec = A;
                    u.Y = Y;
                    u.Jn = J3;  // x_{k-2} is dead once Y is formed
                    u.J4 = J4;
                    u.acc = (k >= 2 && k < numit);
                    u.write_j4 = (k < numit);
                    u.partials = part;
                    u.r2c_out = 0;
                    if (!(fuse_epi && k == numit)) run_x(vp, u, z0, nz);
                });
                std::swap(J2, J3);
                if (dbg) dbg_after(k, J2);
            }
            };
            if (dbg) biggs_loop();
            else run_graphed(loop_key, biggs_loop);
            if (fuse_epi) {  // last update + epilogue
                XArgs u{};
                u.mode = XM_BIGGS_FINAL;
                u.spec = A;
                u.Y = Y;
                u.ep = e;
                run_x(vp, u);
            }
            J = (numit % 2 == 0) ? ws.J.as<float>() : ws.J3.as<float>();  // x_numit after the swaps
        }
    }

    if (!fuse_epi) {
        e.J = (method == 1 && numit == 0 && !(o->flags & PK_FLAG_UNFUSED)) ? I : J;  // OMW with no iteration: J = I
        hipEvent_t ee = prof_mark();
        hipLaunchKernelGGL(k_epilogue, dim3(epi_grid(e)), dim3(256), 0, stream, e);
        PK_HIP(hipGetLastError());
        prof_done("epilogue", (double)on * (4.0 + (e.damp > 1.f ? 4.0 : 0.0) + (double)osz), ee);
    }
    if (out->mem == PK_MEM_HOST)
        PK_HIP(hipMemcpyAsync(out->data, dout, osz * on, hipMemcpyDeviceToHost, stream));
    PK_HIP(hipStreamSynchronize(stream));
    prof_flush();
    if (err_mat) {
        std::memset(err_mat, 0, sizeof(double) * 4 * (size_t)numit);
        if (dbg && (int)mse.size() == numit) debug_err_mat(mse, dbg_step, err_mat);
    }
}

// err_mat of decon_lucy_function.m:144-157 / decon_lucy_omw_function.m:114-127
// (NUMIT x 4, column-major) from the per-iteration MSE, row for row as the
// reference fills it -- including its use of row istp (not iteration istp *
// estep) in the criterion, and of the latest istp for later rows.
void debug_err_mat(const std::vector<double>& mse, int estep, double* e) {
    const int n = (int)mse.size();
    auto E = [&](int row, int col) -> double& { return e[(size_t)(row - 1) + (size_t)n * (size_t)(col - 1)]; };
    int istp = 0;
    for (int k = 1; k <= n; ++k) {
        E(k, 1) = k;
        E(k, 2) = mse[(size_t)k - 1];
        E(k, 3) = E(k, 2) / E(1, 2);
        if (k % estep == 0 && k > estep * 2) {
            istp = k / estep;
            E(k, 4) = std::min(std::fabs(E(istp, 3) - E(istp - 1, 3)) / 10,
                               std::fabs(E(istp, 3) + E(istp - 2, 3) - 2 * E(istp - 1, 3)) / 2);
        } else if (k < 3 * estep) {
            E(k, 4) = 1;
        } else {
            E(k, 4) = E(istp * estep, 4);
        }
    }
}

// ------------------------------------------------------------------ whole-volume edge erosion
void pk_ctx_impl::edge_erosion_mask(const pk_vol* im, int E, pk_vol* mask) {
    if (!im || !mask || !im->data || !mask->data) fail(PK_ERR_INVALID_ARG, "null volume");
    if (im->dtype != PK_UINT16 && im->dtype != PK_SINGLE) fail(PK_ERR_UNSUPPORTED_TYPE, "im must be uint16 or single");
    if (mask->dtype != PK_UINT8) fail(PK_ERR_INVALID_ARG, "the mask is PK_UINT8");
    if (E < 1 || E > 256) fail(PK_ERR_INVALID_ARG, "EdgeErosion must lie in [1, 256]");
    for (int d = 0; d < 3; ++d)
        if (im->dims[d] < 1 || mask->dims[d] != im->dims[d]) fail(PK_ERR_INVALID_ARG, "mask dims must equal im dims");
    const int64_t n1 = im->dims[0], n2 = im->dims[1], n3 = im->dims[2], N = n1 * n2 * n3;
    const size_t esz = im->dtype == PK_UINT16 ? 2 : 4;
    const void* raw = im->data;
    if (im->mem == PK_MEM_HOST) {
        ensure(ws.raw, esz * (size_t)N);
        PK_HIP(hipMemcpyAsync(ws.raw.p, im->data, esz * (size_t)N, hipMemcpyHostToDevice, stream));
        raw = ws.raw.p;
    }
    const size_t dbytes = (sizeof(uint16_t) * (size_t)N + 255) / 256 * 256;
    ensure(ws.tmp, 2 * dbytes + (mask->mem == PK_MEM_HOST ? (size_t)N : 0));
    uint16_t* d1 = ws.tmp.as<uint16_t>();
    uint16_t* d2 = reinterpret_cast<uint16_t*>(ws.tmp.as<char>() + dbytes);
    uint8_t* m = mask->mem == PK_MEM_HOST ? reinterpret_cast<uint8_t*>(ws.tmp.as<char>() + 2 * dbytes)
                                          : static_cast<uint8_t*>(mask->data);
    const int R = E - 1;
    const uint16_t* r16 = im->dtype == PK_UINT16 ? static_cast<const uint16_t*>(raw) : nullptr;
    const float* rf = im->dtype == PK_SINGLE ? static_cast<const float*>(raw) : nullptr;
    hipLaunchKernelGGL(k_edt_mask_pass1, dim3(grid_for(N)), dim3(256), 0, stream, r16, rf, n1, n2, n3, R, d1);
    hipLaunchKernelGGL(k_edt_mask_pass, dim3(grid_for(N)), dim3(256), 0, stream, d1, d2, n1, n2, n3, 2, R);
    hipLaunchKernelGGL(k_edt_mask_pass, dim3(grid_for(N)), dim3(256), 0, stream, d2, d1, n1, n2, n3, 3, R);
    hipLaunchKernelGGL(k_edt_mask_final, dim3(grid_for(N)), dim3(256), 0, stream, d1, m, N, R);
    PK_HIP(hipGetLastError());
    if (mask->mem == PK_MEM_HOST)
        PK_HIP(hipMemcpyAsync(mask->data, m, (size_t)N, hipMemcpyDeviceToHost, stream));
    PK_HIP(hipStreamSynchronize(stream));
}

// ------------------------------------------------------------------ no-allocation OMW
namespace {
// Carves 256-byte aligned pieces off a device region (base may be null to
// measure: the pieces' addresses are then offsets).
struct Carver {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t bytes) {
        off = (off + 255) / 256 * 256;
        T* p = reinterpret_cast<T*>(base + off);
        off += bytes;
        return p;
    }
};

struct ProfOff {  // per-pass events stay off for the call (an event is a resource)
    bool& flag;
    bool was;
    explicit ProfOff(bool& f) : flag(f), was(f) { flag = false; }
    ~ProfOff() { flag = was; }
};

// PSF ./ sum(PSF(:)) with the sum accumulated in double and the division in
// single, as the cached path's otf_build does
std::vector<float> unit_sum(const float* psf, const int64_t pd[3]) {
    const int64_t P = pd[0] * pd[1] * pd[2];
    double s = 0;
    for (int64_t i = 0; i < P; ++i) s += (double)psf[i];
    const float sf = (float)s;
    std::vector<float> q((size_t)P);
    for (int64_t i = 0; i < P; ++i) q[(size_t)i] = psf[i] / sf;
    return q;
}

struct NoallocLayout {
    float2 *tx, *ty, *tz, *rt;
    float *I, *J;
    float2 *A, *Hf, *Hb;
    uint8_t *m1, *m2;
    unsigned long long* cnt;
    size_t bytes;
};

NoallocLayout noalloc_layout(char* base, const int64_t d[3], int edge_erosion) {
    const int64_t n1 = d[0], n2 = d[1], n3 = d[2], N = n1 * n2 * n3;
    const int64_t h1 = n1 / 2 + 1, hp = (h1 + 15) / 16 * 16, S = hp * n2 * n3;
    const bool packed = n1 % 2 == 0;
    Carver c{base};
    NoallocLayout L{};
    L.tx = c.take<float2>(sizeof(float2) * (size_t)(packed ? n1 / 2 : n1));
    L.ty = c.take<float2>(sizeof(float2) * (size_t)n2);
    L.tz = c.take<float2>(sizeof(float2) * (size_t)n3);
    L.rt = packed ? c.take<float2>(sizeof(float2) * (size_t)(n1 / 2 + 1)) : nullptr;
    L.I = c.take<float>(sizeof(float) * (size_t)N);
    L.J = c.take<float>(sizeof(float) * (size_t)N);
    L.A = c.take<float2>(sizeof(float2) * (size_t)S);
    L.Hf = c.take<float2>(sizeof(float2) * (size_t)S);
    L.Hb = c.take<float2>(sizeof(float2) * (size_t)S);
    if (edge_erosion > 0) {
        L.m1 = c.take<uint8_t>((size_t)N);
        L.m2 = c.take<uint8_t>((size_t)N);
        L.cnt = c.take<unsigned long long>(32);
    }
    L.bytes = c.off;
    return L;
}
}  // namespace

size_t pk_ctx_impl::noalloc_bytes(const int64_t d[3], int edge_erosion) {
    return noalloc_layout(nullptr, d, edge_erosion).bytes;
}

void pk_ctx_impl::decon_omw_noalloc(const pk_vol* Iv, const float* psf_f, const int64_t fd[3], const float* psf_b,
                                    const int64_t bd[3], int numit, const pk_opts* o, pk_vol* out, double* err_mat,
                                    void* scratch, size_t scratch_bytes) {
    const DeconGeom g = decon_check(1, Iv, psf_f, fd, psf_b, bd, numit, o, out);
    if (Iv->mem != PK_MEM_DEVICE || out->mem != PK_MEM_DEVICE)
        fail(PK_ERR_INVALID_ARG, "the no-allocation entry point takes device volumes only (I and out)");
    if (o->flags & (PK_FLAG_UNFUSED | PK_FLAG_AUTOTUNE))
        fail(PK_ERR_INVALID_ARG, "PK_FLAG_UNFUSED and PK_FLAG_AUTOTUNE allocate: not accepted without allocation");
    if (!scratch) fail(PK_ERR_HIP, "the device scratch array is not available on this device");
    const int64_t n1 = Iv->dims[0], n2 = Iv->dims[1], n3 = Iv->dims[2];
    const size_t need = noalloc_bytes(Iv->dims, o->edge_erosion);
    if (need > scratch_bytes)
        fail(PK_ERR_UNSUPPORTED_SIZE, "a " + std::to_string(n1) + "x" + std::to_string(n2) + "x" + std::to_string(n3) +
                                          " volume needs " + std::to_string(need) + " bytes of device scratch; the "
                                          "no-allocation entry point holds " + std::to_string(scratch_bytes));
    ProfOff no_events(prof_on);
    VolPlan vp = plan_shapes(n1, n2, n3);  // host only: shapes and pass order, no tables
    const NoallocLayout L = noalloc_layout(static_cast<char*>(scratch), Iv->dims, o->edge_erosion);
    // plan tables into the scratch (the host vectors live until the stream sync below)
    std::vector<float2> tx, ty, tz, rt;
    twiddle_host_tables(vp, tx, ty, tz, rt);
    PK_HIP(hipMemcpyAsync(L.tx, tx.data(), sizeof(float2) * tx.size(), hipMemcpyHostToDevice, stream));
    PK_HIP(hipMemcpyAsync(L.ty, ty.data(), sizeof(float2) * ty.size(), hipMemcpyHostToDevice, stream));
    PK_HIP(hipMemcpyAsync(L.tz, tz.data(), sizeof(float2) * tz.size(), hipMemcpyHostToDevice, stream));
    vp.x.fp.tw = L.tx;
    vp.y.fp.tw = L.ty;
    vp.z.fp.tw = L.tz;
    if (vp.packed) {
        PK_HIP(hipMemcpyAsync(L.rt, rt.data(), sizeof(float2) * rt.size(), hipMemcpyHostToDevice, stream));
        vp.rtw = L.rt;
    }
    // OTF_f and OTF_b (decon_lucy_omw_function.m:70-71) built in place: J (not
    // yet live) is the zero-padded PSF volume, A (not yet live) holds the
    // uploaded PSF values
    const std::vector<float> qf = unit_sum(psf_f, fd), qb = unit_sum(psf_b, bd);
    build_otf_into(vp, qf.data(), fd, L.Hf, L.J, reinterpret_cast<float*>(L.A));
    build_otf_into(vp, qb.data(), bd, L.Hb, L.J, reinterpret_cast<float*>(L.A));

    const uint16_t* raw16 = Iv->dtype == PK_UINT16 ? static_cast<const uint16_t*>(Iv->data) : nullptr;
    const float* rawf = Iv->dtype == PK_SINGLE ? static_cast<const float*>(Iv->data) : nullptr;
    const uint8_t* mask = nullptr;
    if (o->edge_erosion > 0) mask = edge_mask(raw16, rawf, n1, n2, n3, (int)o->edge_erosion, L.m1, L.m2, L.cnt);
    EpiArgs e = epi_args(vp, o, g, L.I, mask, raw16, rawf, out->data);
    const bool fuse_epi = !(o->flags & PK_FLAG_EPILOGUE_PASS) && numit > 0;
    const XArgs pa = prep_args(1, L.A, L.I, L.J, raw16, rawf, o->background);
    omw_passes(vp, pa, L.I, L.J, L.A, L.Hf, L.Hb, numit, e, fuse_epi, 0, nullptr);
    if (!fuse_epi) {
        e.J = numit == 0 ? L.I : L.J;  // OMW with no iteration: J = I
        hipLaunchKernelGGL(k_epilogue, dim3(epi_grid(e)), dim3(256), 0, stream, e);
        PK_HIP(hipGetLastError());
    }
    PK_HIP(hipStreamSynchronize(stream));
    if (err_mat) std::memset(err_mat, 0, sizeof(double) * 4 * (size_t)numit);
}

}  // namespace pk
