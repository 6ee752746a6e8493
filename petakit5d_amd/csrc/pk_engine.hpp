// pk_engine.hpp -- host side of the deconvolution engine: transform plans,
// device workspace, the OTF cache and the RL / OMW iteration drivers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pkdecon.h"
#include "pk_fft_dev.hpp"

namespace pk {

struct Error {
    pk_status st;
    std::string msg;
};

[[noreturn]] void fail(pk_status st, const std::string& msg);
void hip_check(hipError_t e, const char* what);
#define PK_HIP(x) ::pk::hip_check((x), #x)

bool smooth235(int64_t n);

// OTF cache keys: content hash of a PSF and its size; key_mix folds further
// integers (a crop window) into a key
uint64_t psf_key(const float* psf, const int64_t pd[3]);
uint64_t key_mix(uint64_t h, const int64_t* v, int n);

// Device allocation owned by a context (counted against its budget).
struct pk_ctx_impl;
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    pk_ctx_impl* owner = nullptr;
    bool pooled = false;  // from the context's block cache: release returns it there
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf();
    void release();
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct AxisLaunch {
    FftPlan fp{};
    int tile = 1;      // rows (x) or columns (y, z) per block
    int lpad = 0;      // x: LDS row pitch
    dim3 block{1, 1, 1};
    size_t lds = 0;
    int nt = 256;      // y, z: block size the kernel is compiled for (256, 512 or 1024)
};

struct VolPlan {
    // S: float2 elements of a (pitched) spectrum buffer; Sa = h1*n2*n3, the
    // algorithmic half-spectrum size used for byte accounting
    int64_t n1 = 0, n2 = 0, n3 = 0, h1 = 0, hp = 0, N = 0, S = 0, Sa = 0;
    int packed = 0;
    AxisLaunch x, y, z;
    const float2* rtw = nullptr;  // W_{n1}^k, k in [0, n1/2]
    // pass order: the strided axis (2 or 3) that carries the fused forward
    // transform x OTF x inverse transform; the other strided axis takes a
    // forward pass before it and an inverse pass after it
    int conv_axis = 3;
    bool tuned = false;  // launch shapes and pass order measured by autotune()
    // complex points of a row transform: n1 / 2 for packed (even) rows, n1 for odd rows
    int row_len() const { return (int)(packed ? n1 / 2 : n1); }
};

// Host twiddle tables of a plan: its row length, n2, n3 and (even n1) the
// split step's W_{n1}^k, k in [0, n1/2] -- the contents of the device tables
// plan() uploads.
void twiddle_host_tables(const VolPlan& vp, std::vector<float2>& tx, std::vector<float2>& ty,
                         std::vector<float2>& tz, std::vector<float2>& rt);

struct OtfEntry {
    int64_t dims[3];
    int64_t pdims[3];
    uint64_t hash;
    uint64_t tick;
    int pins = 0;  // volumes in flight that read this OTF: never evicted while > 0
    std::unique_ptr<DevBuf> spec;
};

// One volume in flight in the step engine (batched or served).
struct Chunk {
    int method = 1, numit = 0, k = 0, par = 0;  // k: iterations done; par: Biggs buffer parity
    pk_opts o{};
    const VolPlan* vp = nullptr;
    int64_t d[3] = {0, 0, 0};  // caller's size (the plan may be padded)
    const float2* Hf = nullptr;
    const float2* Hb = nullptr;
    float *I = nullptr, *J = nullptr, *J3 = nullptr, *Y = nullptr, *J4 = nullptr, *lam = nullptr;
    float2* A = nullptr;
    double* part = nullptr;
    const uint16_t* raw16 = nullptr;
    const float* rawf = nullptr;
    const void* raw_user = nullptr;
    bool u16 = true, in_host = false;
    void* dout = nullptr;
    void* out_host = nullptr;
    size_t out_bytes = 0;
    bool done() const;
};

// BackProjector.m (host): see pk_wb_backprojector in pkdecon.h
void wb_backprojector(const float* psf, const int64_t d[3], int bp_type, double alpha, double beta, int n,
                      int res_flag, const double ires[3], double* psf_bp, double* otf_bp, double* info);

// Output geometry of a decon call: 0-based bbox start and output size.
struct DeconGeom {
    int64_t b0[3], od[3];
};
// Argument checks of pk_decon_lucy / pk_decon_lucy_omw (method 0 / 1).
DeconGeom decon_check(int method, const pk_vol* I, const float* psf_f, const int64_t fd[3], const float* psf_b,
                      const int64_t bd[3], int numit, const pk_opts* o, const pk_vol* out);
// Epilogue arguments (decon_lucy_function.m:177-203) writing to dout.
EpiArgs epi_args(const struct VolPlan& vp, const pk_opts* o, const DeconGeom& g, const float* I,
                 const uint8_t* mask, const uint16_t* raw16, const float* rawf, void* dout);

// outSize [ny nx' nz'] of deskewRotateFrame3D (deskewRotateFrame3D.m:61-69;
// with resampleFactor rs (x y z order, null = none) round(outSize ./ rs([2 1 3])), :185-191)
void dsr_output_size(const int64_t dims[3], double angle, double dz, double xy, bool reverse, bool objective,
                     double x_step_thresh, int64_t out[3], const double* rs = nullptr);

// err_mat of the reference's debug branch from the per-iteration MSE (NUMIT x 4, column-major)
void debug_err_mat(const std::vector<double>& mse, int estep, double* err_mat);

// Resolves the no-allocation entry point's device scratch array on `device`
// (the current device); called once per context creation (pk_noalloc.hip).
void noalloc_setup(int device);

struct Server;
// request loop (pk_serve.hip)
void serve_start(struct pk_ctx_impl* ctx, int max_active);
void serve_stop(struct pk_ctx_impl* ctx);
uint64_t serve_submit(struct pk_ctx_impl* ctx, const pk_request& q);
void serve_poll(struct pk_ctx_impl* ctx, uint64_t id, pk_req_info* info, double wait_ms, std::string* err);
void serve_cancel(struct pk_ctx_impl* ctx, uint64_t id);
void serve_release(struct pk_ctx_impl* ctx, uint64_t id);
pk_serve_stats serve_stats(struct pk_ctx_impl* ctx);

// Pinned host staging buffer (hipHostMalloc), so uploads are truly asynchronous.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
};

struct Workspace {
    DevBuf I, J, A, J3, Y, J4, partials, lambda, raw, out, mask, mask2, count, tmp;
    DevBuf batch;  // batched decon: per-chunk arena
    DevBuf dsr;    // deskew-rotate: skewed-space interpolated volume / conversions
    DevBuf dbg;    // debug metrics: block partials of the squared differences and their sum
    // per-step descriptor tables, double-buffered: step s stages its table in
    // tab_host[s&1] and copies it to tab_dev[s&1]; tab_ev[s&1] marks the copy's
    // completion, so the host can build step s+1 while step s runs
    DevBuf tab_dev[2];
    HostBuf tab_host[2];
    hipEvent_t tab_ev[2] = {nullptr, nullptr};
    int tab_slot = 0;
};

// Per-pass timing with HIP events recorded on the engine's stream.
struct ProfAgg {
    int64_t launches = 0;
    double total_ms = 0;
    double bytes = 0;  // algorithmic bytes per launch
};
struct ProfPending {
    std::string name;
    double bytes;
    hipEvent_t e0, e1;
};

struct pk_ctx_impl {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t live_bytes = 0;
    std::recursive_mutex mu;
    std::string last_error;
    std::map<int, std::unique_ptr<DevBuf>> tw;    // per transform length
    std::map<int64_t, std::unique_ptr<DevBuf>> rtw;  // per even n1
    std::map<std::vector<int64_t>, VolPlan> plans;
    std::vector<OtfEntry> otf;
    uint64_t tick = 0;
    size_t otf_max_entries = 8;
    Workspace ws;
    bool prof_on = false;
    // batched steps: chunk sets run their passes on the engine's stream and up
    // to kSideStreams more (default priority), forked after the step's table
    // upload and joined at its end
    static constexpr int kSideStreams = 7;  // at most 8 sets; chunk_streams() picks how many
    hipStream_t side[kSideStreams] = {};
    hipStream_t copy_stream = nullptr;  // served requests' output copies (created on first use)
    hipStream_t h2d_stream = nullptr;   // served requests' input copies and their arenas (created on first use)
    hipEvent_t ev_copy = nullptr;       // orders a host copy after the engine stream's work
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[kSideStreams] = {};
    std::vector<ProfPending> prof_pending;
    std::vector<hipEvent_t> ev_pool;
    std::map<std::string, ProfAgg> prof;

    hipEvent_t prof_mark();                       // records an event (profiling on) or returns null
    void prof_done(const char* name, double bytes, hipEvent_t e0);
    void prof_flush();                            // after a stream sync: fold event pairs into prof

    // device memory helpers
    void alloc(DevBuf& b, size_t bytes);
    // Device blocks kept for reuse by size class (request arenas of the
    // server): allocation takes a cached block of the size's class when there
    // is one, release puts the block back; nothing synchronises the device.
    // The caller releases a block only when no queued work still uses it.
    // Cached bytes count against the HBM budget; an allocation that would
    // exceed it first frees the cache.
    std::multimap<size_t, void*> block_cache;
    size_t block_cache_bytes = 0;
    void alloc_pooled(DevBuf& b, size_t bytes);
    void block_put(void* p, size_t bytes);
    void block_cache_clear();
    void ensure(DevBuf& b, size_t bytes);

    const VolPlan& plan(int64_t n1, int64_t n2, int64_t n3);
    // geometry, launch shapes (built-in or from the tuning table) and pass
    // order of a padded size, without device tables (their pointers null)
    VolPlan plan_shapes(int64_t n1, int64_t n2, int64_t n3) const;
    const float2* otf_for(const VolPlan& vp, const float* psf, const int64_t pdims[3]);
    // cache lookup by key (nullptr on a miss) and build of a missing entry
    const float2* otf_find(const VolPlan& vp, uint64_t key, const int64_t pdims[3]);
    const float2* otf_build(const VolPlan& vp, uint64_t key, const float* psf, const int64_t pdims[3]);
    // R2C of the zero-padded, circularly shifted PSF q (decon_psf2otf.m:19-27) into dst
    void build_otf(const VolPlan& vp, const float* q, const int64_t pdims[3], float2* dst);
    // the same on the engine's stream with the caller's device buffers: real
    // (N floats, zeroed here) and dq (the PSF's prod(pdims) floats, uploaded
    // here from q, which must stay valid until the stream reaches the copy)
    void build_otf_into(const VolPlan& vp, const float* q, const int64_t pdims[3], float2* dst, float* real,
                        float* dq);

    // transforms (device pointers)
    void r2c3d(const VolPlan& vp, const float* real, float2* spec);
    void c2r3d(const VolPlan& vp, float2* spec, float* real);
    // z0, nz: the planes [z0, z0 + nz) of dim 3 (a slab); nz = 0 means all
    void run_x(const VolPlan& vp, XArgs a, int64_t z0 = 0, int64_t nz = 0);
    void run_col(const VolPlan& vp, int axis, float2* data, int mode, const float2* otf, float scale, int64_t z0 = 0,
                 int64_t nz = 0);
    int64_t slab_planes(const VolPlan& vp) const;
    int chunk_streams() const;  // streams a batched step may use (PK_CHUNK_STREAMS, default 4)
    std::string autotune(int64_t n1, int64_t n2, int64_t n3);
    void set_shapes(int64_t n1, int64_t n2, int64_t n3, const std::array<int, 8>& v);
    bool apply_shapes(VolPlan& vp, const std::array<int, 8>& v);
    // Iteration loops replayed as HIP graphs: a key's first call runs eagerly,
    // the second captures and instantiates, later calls launch the graph.
    struct GraphEntry {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
        uint64_t tick = 0;
    };
    std::vector<GraphEntry> graphs;
    uint64_t graph_tick = 0;
    template <class Fn>
    void run_graphed(const std::vector<uint64_t>& key, Fn&& fn);
    void clear_graphs();
    // convolution of the spectrum in ws.A: forward y, z-conv, inverse y

    void decon(int method, const pk_vol* I, const float* psf_f, const int64_t fd[3], const float* psf_b,
               const int64_t bd[3], int numit, const pk_opts* o, pk_vol* out, double* err_mat);
    // decon_lucy_omw_function with every device buffer (plan tables, OTFs,
    // estimate, spectrum, masks) carved from the caller's device region
    // [scratch, scratch + scratch_bytes): no allocation, no cache, no graph.
    // I and out must be device memory.
    void decon_omw_noalloc(const pk_vol* I, const float* psf_f, const int64_t fd[3], const float* psf_b,
                           const int64_t bd[3], int numit, const pk_opts* o, pk_vol* out, double* err_mat,
                           void* scratch, size_t scratch_bytes);
    // bytes of scratch decon_omw_noalloc needs for a volume of size d
    static size_t noalloc_bytes(const int64_t d[3], int edge_erosion);
    // the OMW pass sequence shared by decon and decon_omw_noalloc
    // after (optional): called with k and the estimate after every iteration
    void omw_passes(const VolPlan& vp, const XArgs& prep, float* I, float* J, float2* A, const float2* Hf,
                    const float2* Hb, int numit, const EpiArgs& e, bool fuse_epi, int64_t kz,
                    const std::vector<uint64_t>* loop_key,
                    const std::function<void(int, const float*)>* after = nullptr, uint16_t* I16 = nullptr);
    // the plan's row passes run the compile-time kernel (the XM_*_U16 modes need it)
    bool x_compile_time(const VolPlan& vp) const;
    // debug ('debug', true): per-iteration MSE and the iteration hook
    int32_t dbg_step = 1;
    pk_iter_hook dbg_hook = nullptr;
    void* dbg_user = nullptr;
    // decon_mask_edge_erosion of I ~= 0 into the caller's buffers; returns the eroded mask
    const uint8_t* edge_mask(const uint16_t* raw16, const float* rawf, int64_t n1, int64_t n2, int64_t n3, int E,
                             uint8_t* m1, uint8_t* m2, unsigned long long* cnt);
    void reset();
    // deskewRotateFrame3D (pk_dsr.hip); bbox: 1-based inclusive output box
    // or null; device volumes, returns when done; interp: 0 linear, 1 cubic
    // (imwarp's form, pk_deskew.hip); rs: resampleFactor (x y z) or null;
    // imwarp_form: the imwarp form even for linear without resampling (the
    // reference's gpuProcess = true)
    void deskew_rotate(const pk_vol* in, double angle, double dz, double xy, bool reverse, bool objective,
                       double x_step_thresh, const int64_t* bbox, pk_vol* out, int interp = 0,
                       const double* rs = nullptr, bool imwarp_form = false);
    // bwdist(~bw) > E - 1 of bw = im > 0 with cleared faces (RLdecon_large_in_memory.m:150-155)
    void edge_erosion_mask(const pk_vol* im, int E, pk_vol* mask);
    // copy of a pitched half spectrum with one pass's access pattern (pk_probe.hip): ms per copy
    // (mean of reps) and the bytes it moves
    // flags: bit 0 XCD-aware order, bit 1 cold copies (memory-side cache flushed before each)
    double probe_copy(int pattern, const int64_t dims[3], int tile, int vals, int flags, int reps, double* bytes);
    void decon_batch(int method, const pk_batch_item* items, int n, const float* psf_f, const int64_t fd[3],
                     const float* psf_b, const int64_t bd[3], int numit, const pk_opts* o);
    // step engine
    void chunk_validate(const pk_vol& in, const pk_vol& out, const pk_opts& o);
    // hf, hb: psf_key of psf_f / psf_b (hashed once per batch or request)
    size_t chunk_prepare(Chunk& c, const pk_vol& in, const pk_vol& out, const float* psf_f, const int64_t fd[3],
                         uint64_t hf, const float* psf_b, const int64_t bd[3], uint64_t hb);
    // copy_stream (optional): the stream of the chunk's host copy, so it
    // overlaps the steps of other chunks.  chunk_place: the arena at base must
    // have no queued work on other streams (a cached block whose users have
    // completed, or fresh memory); chunk_finish: the copy waits for the work
    // queued on the engine stream.
    void chunk_place(Chunk& c, char* base, hipStream_t copy_stream = nullptr);
    void run_step(std::vector<Chunk*>& act);
    void chunk_finish(Chunk& c, hipStream_t copy_stream = nullptr);
    void copy_after_stream(hipStream_t copy_stream);  // copy_stream waits for the work queued on stream so far
    void otf_pin(const float2* spec, int delta);
    // MATLAB-structured comparison path (pk_plain.hip); leaves the estimate in J
    void decon_unfused(int method, const VolPlan& vp, const void* raw, bool u16, float bg, const float2* Hf,
                       const float2* Hb, int numit, float* I, float* J);
    std::unique_ptr<Server> srv;
    pk_ctx_impl();  // defined where Server is complete (pk_serve.hip)
    ~pk_ctx_impl();

    // omw_backprojector_generation (host steps + GPU nearest-edge search).
    // thr = {OTFCumThresh, OTFAreaThresh, hann lo, hann hi, minIntThrsh}.
    void omw_backprojector(const float* psf, const int64_t pd[3], double alpha, int skewed, const double thr[5],
                           float* b_omw, float* otf_bp_omw, float* abs_otf_c, float* otf_mask, int* skewed_used);
};

}  // namespace pk
