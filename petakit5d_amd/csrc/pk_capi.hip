// pk_capi.hip -- the extern "C" boundary (include/pkdecon.h).  Every entry
// point validates its arguments, serialises on the context, converts engine
// exceptions into a pk_status plus pk_last_error() text, and never aborts.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/pkdecon.h"
#include "pk_engine.hpp"

struct pk_ctx : pk::pk_ctx_impl {};

namespace pk {
struct DeskewGeom {
    int64_t nx_out;
    double xstep, xshift;
};
DeskewGeom deskew_geometry(const int64_t dims[3], double angle, double dz, double xy, bool reverse, bool crop);
void launch_deskew(const float* vol, float* out, const int64_t dims[3], const DeskewGeom& g, hipStream_t s);
struct Affine3 {
    double a[4][4];
};
void launch_affine3d(const float* vol, float* out, const int64_t dims[3], const int64_t odims[3], const Affine3& m,
                     hipStream_t s);
void launch_affine3d_cubic(const float* vol, float* out, const int64_t dims[3], const int64_t odims[3],
                           const Affine3& m, hipStream_t s);
}  // namespace pk

namespace {
template <class Fn>
pk_status guarded(pk_ctx* c, Fn&& fn) {
    if (!c) return PK_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    try {
        PK_HIP(hipSetDevice(c->device));
        fn();
        c->last_error.clear();
        return PK_OK;
    } catch (const pk::Error& e) {
        c->last_error = e.msg;
        return e.st;
    } catch (const std::bad_alloc&) {
        c->last_error = "host allocation failed";
        return PK_ERR_NO_MEMORY;
    } catch (...) {
        c->last_error = "unexpected exception";
        return PK_ERR_INVALID_ARG;
    }
}
}  // namespace

// Serving entry points take the request-queue lock only, never the context
// lock the scheduler holds while it runs a step.
template <class Fn>
static pk_status serve_guarded(pk_ctx* c, Fn&& fn) {
    if (!c) return PK_ERR_INVALID_ARG;
    try {
        fn();
        return PK_OK;
    } catch (const pk::Error& e) {
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        c->last_error = e.msg;
        return e.st;
    } catch (...) {
        return PK_ERR_INVALID_ARG;
    }
}

// Host-only entry points: no device, no context lock; the context (optional)
// only receives the error text.
template <class Fn>
static pk_status host_guarded(pk_ctx* c, Fn&& fn) {
    try {
        fn();
        if (c) {
            std::lock_guard<std::recursive_mutex> lk(c->mu);
            c->last_error.clear();
        }
        return PK_OK;
    } catch (const pk::Error& e) {
        if (c) {
            std::lock_guard<std::recursive_mutex> lk(c->mu);
            c->last_error = e.msg;
        }
        return e.st;
    } catch (const std::bad_alloc&) {
        return PK_ERR_NO_MEMORY;
    } catch (...) {
        return PK_ERR_INVALID_ARG;
    }
}

extern "C" {

void pk_opts_default(pk_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->background = 0.0;
    o->damp_factor = 1.0;
    o->scale_factor = 1.0;
    o->decon_offset = 0.0;
    o->edge_erosion = 0;
    o->save16bit = 1;
    o->has_bbox = 0;
}

pk_status pk_ctx_create(int device, uint64_t hbm_budget_bytes, pk_ctx** out) {
    if (!out) return PK_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return PK_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) return PK_ERR_INVALID_ARG;
    pk_ctx* c = new (std::nothrow) pk_ctx();
    if (!c) return PK_ERR_NO_MEMORY;
    c->device = device;
    c->budget = hbm_budget_bytes;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return PK_ERR_HIP;
    }
    c->stream = c->own_stream;
    pk::noalloc_setup(device);  // the no-allocation entry point's scratch array on this device
    *out = c;
    return PK_OK;
}

void pk_ctx_destroy(pk_ctx* c) {
    if (!c) return;
    pk::serve_stop(c);  // joins the scheduler (which takes the context lock itself)
    {
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        try {
            c->reset();
        } catch (...) {
        }
        for (int i = 0; i < pk_ctx::kSideStreams; ++i) {
            if (c->side[i]) {
                (void)hipStreamSynchronize(c->side[i]);
                (void)hipStreamDestroy(c->side[i]);
            }
            if (c->ev_join[i]) (void)hipEventDestroy(c->ev_join[i]);
        }
        if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
        if (c->ev_copy) (void)hipEventDestroy(c->ev_copy);
        for (hipStream_t* cs : {&c->copy_stream, &c->h2d_stream})
            if (*cs) {
                (void)hipStreamSynchronize(*cs);
                (void)hipStreamDestroy(*cs);
            }
        if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    }
    delete c;
}

pk_status pk_reset(pk_ctx* c) {
    return guarded(c, [&] { c->reset(); });
}

pk_status pk_set_stream(pk_ctx* c, void* s) {
    return guarded(c, [&] {
        PK_HIP(hipStreamSynchronize(c->stream));
        c->stream = s ? static_cast<hipStream_t>(s) : c->own_stream;
    });
}

const char* pk_last_error(const pk_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

uint64_t pk_device_bytes(const pk_ctx* c) { return c ? c->live_bytes : 0; }

pk_status pk_decon_lucy(pk_ctx* c, const pk_vol* I, const float* psf, const int64_t psf_dims[3], int32_t numit,
                        const pk_opts* opts, pk_vol* out, double* err_mat) {
    return guarded(c, [&] {
        if (!psf_dims) pk::fail(PK_ERR_INVALID_ARG, "psf_dims is null");
        c->decon(0, I, psf, psf_dims, nullptr, nullptr, numit, opts, out, err_mat);
    });
}

pk_status pk_decon_lucy_omw(pk_ctx* c, const pk_vol* I, const float* psf_f, const int64_t f_dims[3],
                            const float* psf_b, const int64_t b_dims[3], int32_t numit, const pk_opts* opts,
                            pk_vol* out, double* err_mat) {
    return guarded(c, [&] {
        if (!f_dims || !b_dims) pk::fail(PK_ERR_INVALID_ARG, "PSF dims are null");
        c->decon(1, I, psf_f, f_dims, psf_b, b_dims, numit, opts, out, err_mat);
    });
}

pk_status pk_decon_batch(pk_ctx* c, int32_t method, const pk_batch_item* items, int32_t n, const float* psf_f,
                         const int64_t f_dims[3], const float* psf_b, const int64_t b_dims[3], int32_t numit,
                         const pk_opts* opts) {
    return guarded(c, [&] {
        if (!f_dims || (method == 1 && !b_dims)) pk::fail(PK_ERR_INVALID_ARG, "PSF dims are null");
        if (method != 0 && method != 1) pk::fail(PK_ERR_INVALID_ARG, "method must be 0 (RL) or 1 (OMW)");
        c->decon_batch(method, items, n, psf_f, f_dims, psf_b, b_dims, numit, opts);
    });
}

pk_status pk_psf2otf(pk_ctx* c, const float* psf, const int64_t pd[3], const int64_t od[3], float* otf) {
    return guarded(c, [&] {
        if (!psf || !pd || !od || !otf) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        for (int d = 0; d < 3; ++d) {
            if (pd[d] < 1 || od[d] < 1) pk::fail(PK_ERR_INVALID_ARG, "empty size");
            if (od[d] < pd[d]) pk::fail(PK_ERR_OUT_SIZE, "outSizeIsSmallerThanPsfSize");
        }
        const pk::VolPlan& vp = c->plan(od[0], od[1], od[2]);
        pk::DevBuf spec;
        c->alloc(spec, sizeof(float2) * (size_t)vp.S);
        c->build_otf(vp, psf, pd, spec.as<float2>());
        std::vector<float2> h((size_t)vp.S);
        PK_HIP(hipMemcpy(h.data(), spec.p, sizeof(float2) * (size_t)vp.S, hipMemcpyDeviceToHost));
        // expand the half spectrum with Hermitian symmetry into MATLAB's full OTF
        const int64_t n1 = vp.n1, n2 = vp.n2, n3 = vp.n3, h1 = vp.h1, hp = vp.hp;
        for (int64_t k3 = 0; k3 < n3; ++k3)
            for (int64_t k2 = 0; k2 < n2; ++k2)
                for (int64_t k1 = 0; k1 < n1; ++k1) {
                    float2 v;
                    if (k1 < h1) {
                        v = h[(size_t)(k1 + hp * (k2 + n2 * k3))];
                    } else {
                        const int64_t m1 = n1 - k1, m2 = (n2 - k2) % n2, m3 = (n3 - k3) % n3;
                        v = h[(size_t)(m1 + hp * (m2 + n2 * m3))];
                        v.y = -v.y;
                    }
                    const size_t o = (size_t)(k1 + n1 * (k2 + n2 * k3));
                    otf[2 * o] = v.x;
                    otf[2 * o + 1] = v.y;
                }
    });
}

void pk_bp_opts_default(pk_bp_opts* o) {
    if (!o) return;
    o->otf_cum_thresh = 0.9;
    o->otf_area_thresh = 100;
    o->hann_win_lo = 0.8;
    o->hann_win_hi = 1.0;
    o->min_int_thresh = 1e-3;
}

pk_status pk_omw_backprojector(pk_ctx* c, const float* psf, const int64_t psf_dims[3], double alpha, int32_t skewed,
                               const pk_bp_opts* opts, float* b_omw, float* otf_bp_omw, float* abs_otf_c,
                               float* otf_mask, int32_t* skewed_used) {
    return guarded(c, [&] {
        if (!psf || !psf_dims || !b_omw) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        for (int d = 0; d < 3; ++d)
            if (psf_dims[d] < 1) pk::fail(PK_ERR_EMPTY_PSF, "A valid nonzero PSF needs to be provided!");
        pk_bp_opts o;
        pk_bp_opts_default(&o);
        if (opts) o = *opts;
        if (!(o.hann_win_hi > o.hann_win_lo)) pk::fail(PK_ERR_INVALID_ARG, "hannWinBounds must be increasing");
        const double thr[5] = {o.otf_cum_thresh, o.otf_area_thresh, o.hann_win_lo, o.hann_win_hi, o.min_int_thresh};
        int used = 0;
        c->omw_backprojector(psf, psf_dims, alpha, skewed, thr, b_omw, otf_bp_omw, abs_otf_c, otf_mask, &used);
        if (skewed_used) *skewed_used = used;
    });
}

pk_status pk_fft_r2c(pk_ctx* c, const float* d_real, float* d_spec, const int64_t dims[3]) {
    return guarded(c, [&] {
        if (!d_real || !d_spec || !dims) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        const pk::VolPlan& vp = c->plan(dims[0], dims[1], dims[2]);
        c->ensure(c->ws.A, sizeof(float2) * (size_t)vp.S);
        c->r2c3d(vp, d_real, c->ws.A.as<float2>());
        // pitched (hp) working spectrum -> the caller's dense (h1) layout
        PK_HIP(hipMemcpy2DAsync(d_spec, sizeof(float2) * vp.h1, c->ws.A.p, sizeof(float2) * vp.hp,
                                sizeof(float2) * vp.h1, (size_t)(vp.n2 * vp.n3), hipMemcpyDeviceToDevice, c->stream));
        PK_HIP(hipStreamSynchronize(c->stream));
    });
}

pk_status pk_fft_c2r(pk_ctx* c, const float* d_spec, float* d_real, const int64_t dims[3]) {
    return guarded(c, [&] {
        if (!d_real || !d_spec || !dims) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        const pk::VolPlan& vp = c->plan(dims[0], dims[1], dims[2]);
        c->ensure(c->ws.A, sizeof(float2) * (size_t)vp.S);
        PK_HIP(hipMemcpy2DAsync(c->ws.A.p, sizeof(float2) * vp.hp, d_spec, sizeof(float2) * vp.h1,
                                sizeof(float2) * vp.h1, (size_t)(vp.n2 * vp.n3), hipMemcpyDeviceToDevice, c->stream));
        c->c2r3d(vp, c->ws.A.as<float2>(), d_real);
        PK_HIP(hipStreamSynchronize(c->stream));
    });
}

static void deskew_check(const int64_t dims[3], double dz, double xy) {
    if (!dims || dims[0] < 1 || dims[1] < 1 || dims[2] < 1) pk::fail(PK_ERR_INVALID_ARG, "volume dimensions must be >= 1");
    if (!(dz >= 0) || !(xy > 0)) pk::fail(PK_ERR_INVALID_ARG, "dz must be >= 0 and xyPixelSize > 0");
}

pk_status pk_deskew_out_dims(const int64_t dims[3], double angle, double dz, double xy_pixel_size, int32_t reverse,
                             int32_t crop, int64_t out_dims[3]) {
    return host_guarded(nullptr, [&] {
        deskew_check(dims, dz, xy_pixel_size);
        if (!out_dims) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        const pk::DeskewGeom g = pk::deskew_geometry(dims, angle, dz, xy_pixel_size, reverse != 0, crop != 0);
        if (g.nx_out < 1) pk::fail(PK_ERR_INVALID_ARG, "deskewed width below one column");
        out_dims[0] = dims[0];
        out_dims[1] = g.nx_out;
        out_dims[2] = dims[2];
    });
}

pk_status pk_deskew_frame3d(pk_ctx* c, const float* d_vol, const int64_t dims[3], double angle, double dz,
                            double xy_pixel_size, int32_t reverse, int32_t crop, float* d_out) {
    return guarded(c, [&] {
        deskew_check(dims, dz, xy_pixel_size);
        if (!d_vol || !d_out) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        const pk::DeskewGeom g = pk::deskew_geometry(dims, angle, dz, xy_pixel_size, reverse != 0, crop != 0);
        if (g.nx_out < 1) pk::fail(PK_ERR_INVALID_ARG, "deskewed width below one column");
        pk::launch_deskew(d_vol, d_out, dims, g, c->stream);
        PK_HIP(hipStreamSynchronize(c->stream));
    });
}

pk_status pk_deskew_rotate_out_dims(const int64_t dims[3], double angle, double dz, double xy_pixel_size,
                                    int32_t reverse, int32_t objective_scan, double x_step_thresh, int64_t out_dims[3]) {
    return host_guarded(nullptr, [&] {
        if (!dims || !out_dims) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        for (int i = 0; i < 3; ++i)
            if (dims[i] < 2) pk::fail(PK_ERR_INVALID_ARG, "every axis needs at least 2 voxels");
        pk::dsr_output_size(dims, angle, dz, xy_pixel_size, reverse != 0, objective_scan != 0, x_step_thresh, out_dims);
        if (out_dims[1] < 1 || out_dims[2] < 1) pk::fail(PK_ERR_INVALID_ARG, "empty output box");
    });
}

pk_status pk_deskew_rotate_frame3d(pk_ctx* c, const pk_vol* in, double angle, double dz, double xy_pixel_size,
                                   int32_t reverse, int32_t objective_scan, double x_step_thresh, const int64_t bbox[6],
                                   pk_vol* out) {
    return guarded(c, [&] {
        c->deskew_rotate(in, angle, dz, xy_pixel_size, reverse != 0, objective_scan != 0, x_step_thresh, bbox, out);
    });
}

pk_status pk_deskew_rotate_frame3d_cubic(pk_ctx* c, const pk_vol* in, double angle, double dz, double xy_pixel_size,
                                         int32_t reverse, int32_t objective_scan, double x_step_thresh,
                                         const int64_t bbox[6], pk_vol* out) {
    return guarded(c, [&] {
        c->deskew_rotate(in, angle, dz, xy_pixel_size, reverse != 0, objective_scan != 0, x_step_thresh, bbox, out, 1);
    });
}

pk_status pk_deskew_rotate_resample_out_dims(const int64_t dims[3], double angle, double dz, double xy_pixel_size,
                                             int32_t reverse, int32_t objective_scan, double x_step_thresh,
                                             const double resample[3], int64_t out_dims[3]) {
    return host_guarded(nullptr, [&] {
        if (!dims || !out_dims) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        for (int i = 0; i < 3; ++i)
            if (dims[i] < 2) pk::fail(PK_ERR_INVALID_ARG, "every axis needs at least 2 voxels");
        pk::dsr_output_size(dims, angle, dz, xy_pixel_size, reverse != 0, objective_scan != 0, x_step_thresh, out_dims,
                            resample);
        if (out_dims[1] < 1 || out_dims[2] < 1) pk::fail(PK_ERR_INVALID_ARG, "empty output box");
    });
}

pk_status pk_deskew_rotate_frame3d_resample(pk_ctx* c, const pk_vol* in, double angle, double dz, double xy_pixel_size,
                                            int32_t reverse, int32_t objective_scan, double x_step_thresh,
                                            const double resample[3], int32_t interp, const int64_t bbox[6],
                                            pk_vol* out) {
    return guarded(c, [&] {
        if (interp < 0 || interp > 3) pk::fail(PK_ERR_INVALID_ARG, "interp: 0 linear, 1 cubic, + 2 for the imwarp form");
        c->deskew_rotate(in, angle, dz, xy_pixel_size, reverse != 0, objective_scan != 0, x_step_thresh, bbox, out,
                         interp & 1, resample, (interp & 2) != 0);
    });
}

pk_status pk_set_iter_hook(pk_ctx* c, int32_t save_step, pk_iter_hook hook, void* user) {
    return guarded(c, [&] {
        if (save_step < 1) pk::fail(PK_ERR_INVALID_ARG, "saveStep must be >= 1");
        c->dbg_step = save_step;
        c->dbg_hook = hook;
        c->dbg_user = user;
    });
}

pk_status pk_edge_erosion_mask(pk_ctx* c, const pk_vol* im, int32_t edge_erosion, pk_vol* mask) {
    return guarded(c, [&] { c->edge_erosion_mask(im, edge_erosion, mask); });
}

pk_status pk_affine_warp3d(pk_ctx* c, const float* d_vol, const int64_t dims[3], const double inv_tform[16],
                           const int64_t out_dims[3], float* d_out) {
    return guarded(c, [&] {
        if (!d_vol || !d_out || !dims || !out_dims || !inv_tform) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        for (int i = 0; i < 3; ++i)
            if (dims[i] < 1 || out_dims[i] < 1) pk::fail(PK_ERR_INVALID_ARG, "volume dimensions must be >= 1");
        pk::Affine3 m;
        for (int i = 0; i < 16; ++i) {
            if (!std::isfinite(inv_tform[i])) pk::fail(PK_ERR_INVALID_ARG, "transform entries must be finite");
            m.a[i / 4][i % 4] = inv_tform[i];
        }
        pk::launch_affine3d(d_vol, d_out, dims, out_dims, m, c->stream);
        PK_HIP(hipStreamSynchronize(c->stream));
    });
}

pk_status pk_affine_warp3d_cubic(pk_ctx* c, const float* d_vol, const int64_t dims[3], const double inv_tform[16],
                                 const int64_t out_dims[3], float* d_out) {
    return guarded(c, [&] {
        if (!d_vol || !d_out || !dims || !out_dims || !inv_tform) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        for (int i = 0; i < 3; ++i)
            if (dims[i] < 2 || out_dims[i] < 1)
                pk::fail(PK_ERR_INVALID_ARG, "cubic: input dimensions must be >= 2, output dimensions >= 1");
        pk::Affine3 m;
        for (int i = 0; i < 16; ++i) {
            if (!std::isfinite(inv_tform[i])) pk::fail(PK_ERR_INVALID_ARG, "transform entries must be finite");
            m.a[i / 4][i % 4] = inv_tform[i];
        }
        pk::launch_affine3d_cubic(d_vol, d_out, dims, out_dims, m, c->stream);
        PK_HIP(hipStreamSynchronize(c->stream));
    });
}

pk_status pk_profile_enable(pk_ctx* c, int32_t enable) {
    return guarded(c, [&] {
        PK_HIP(hipStreamSynchronize(c->stream));
        c->prof_flush();
        c->prof.clear();
        c->prof_on = enable != 0;
    });
}

pk_status pk_probe_copy(pk_ctx* c, int32_t pattern, const int64_t dims[3], int32_t tile, int32_t vals_per_thread,
                        int32_t flags, int32_t reps, double* ms, double* bytes) {
    return guarded(c, [&] {
        if (!dims || !ms || !bytes) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        if (pattern < 0 || pattern > 3 || reps < 1) pk::fail(PK_ERR_INVALID_ARG, "pattern 0..3, reps >= 1");
        for (int d = 0; d < 3; ++d)
            if (dims[d] < 2) pk::fail(PK_ERR_INVALID_ARG, "probe dimensions must be >= 2");
        if (flags & ~3) pk::fail(PK_ERR_INVALID_ARG, "probe flags: bit 0 XCD order, bit 1 cold");
        *ms = c->probe_copy(pattern, dims, tile, vals_per_thread, flags, reps, bytes);
    });
}

int32_t pk_profile_read(pk_ctx* c, pk_prof_entry* out, int32_t max) {
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    int32_t n = 0;
    for (const auto& kv : c->prof) {
        if (out && n < max) {
            pk_prof_entry& e = out[n];
            std::memset(&e, 0, sizeof(e));
            std::strncpy(e.name, kv.first.c_str(), sizeof(e.name) - 1);
            e.launches = kv.second.launches;
            e.total_ms = kv.second.total_ms;
            e.bytes_per_launch = kv.second.bytes;
        }
        ++n;
    }
    return n;
}

pk_status pk_serve_start(pk_ctx* c, int32_t max_active) {
    return guarded(c, [&] { pk::serve_start(c, max_active); });
}

pk_status pk_serve_stop(pk_ctx* c) {
    return serve_guarded(c, [&] { pk::serve_stop(c); });
}

pk_status pk_serve_submit(pk_ctx* c, const pk_request* req, uint64_t* id) {
    return serve_guarded(c, [&] {
        if (!req || !id) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        *id = pk::serve_submit(c, *req);
    });
}

pk_status pk_serve_wait(pk_ctx* c, uint64_t id, double wait_ms, pk_req_info* info) {
    return serve_guarded(c, [&] {
        std::string err;
        pk_req_info tmp{};
        pk::serve_poll(c, id, info ? info : &tmp, wait_ms, &err);
        if (!err.empty()) {
            std::lock_guard<std::recursive_mutex> lk(c->mu);
            c->last_error = err;
        }
    });
}

pk_status pk_serve_cancel(pk_ctx* c, uint64_t id) {
    return serve_guarded(c, [&] { pk::serve_cancel(c, id); });
}

pk_status pk_serve_release(pk_ctx* c, uint64_t id) {
    return serve_guarded(c, [&] { pk::serve_release(c, id); });
}

pk_status pk_serve_stats_get(pk_ctx* c, pk_serve_stats* out) {
    return serve_guarded(c, [&] {
        if (!out) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        *out = pk::serve_stats(c);
    });
}

int64_t pk_find_good_factor_number(int64_t n, int32_t dir, int32_t allow_odd_seven, int32_t max_prime) {
    if (dir == 0) dir = 1;
    for (int guard = 0; guard < (1 << 24); ++guard) {
        if (n <= 1) return n;  // factor(1) = 1 and factor(0) = 0 end the search
        int64_t m = n, pmin = 0, pmax = 1;
        for (int64_t p = 2; p * p <= m; ++p)
            while (m % p == 0) {
                if (!pmin) pmin = p;
                pmax = p;
                m /= p;
            }
        if (m > 1) {
            if (!pmin) pmin = m;
            if (m > pmax) pmax = m;
        }
        bool bad = pmax > max_prime;
        if (max_prime >= 7 && !allow_odd_seven && pmin > 2 && pmax == 7) bad = true;
        if (!bad) return n;
        n += dir;
    }
    return -1;
}

pk_status pk_autotune(pk_ctx* c, const int64_t dims[3], char* report, size_t report_len) {
    return guarded(c, [&] {
        if (!dims || dims[0] < 1 || dims[1] < 1 || dims[2] < 1) pk::fail(PK_ERR_INVALID_ARG, "bad dims");
        const std::string r = c->autotune(dims[0], dims[1], dims[2]);
        if (report && report_len) {
            std::strncpy(report, r.c_str(), report_len - 1);
            report[report_len - 1] = 0;
        }
    });
}

pk_status pk_tune_set(pk_ctx* c, const int64_t dims[3], const int32_t shapes[8]) {
    return guarded(c, [&] {
        if (!dims || !shapes || dims[0] < 1 || dims[1] < 1 || dims[2] < 1) pk::fail(PK_ERR_INVALID_ARG, "bad dims");
        std::array<int, 8> v;
        for (int i = 0; i < 8; ++i) v[i] = shapes[i];
        c->set_shapes(dims[0], dims[1], dims[2], v);
    });
}

pk_status pk_wb_backprojector(pk_ctx* c, const float* psf_fp, const int64_t dims[3], int32_t bp_type, double alpha,
                              double beta, int32_t n, int32_t res_flag, const double ires[3], double* psf_bp,
                              double* otf_bp, double* info) {
    return host_guarded(c, [&] {
        if (!psf_fp || !dims || !psf_bp) pk::fail(PK_ERR_INVALID_ARG, "null argument");
        if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1) pk::fail(PK_ERR_INVALID_ARG, "empty PSF");
        if (res_flag == 2 && !ires) pk::fail(PK_ERR_INVALID_ARG, "resFlag 2 needs iRes");
        const double zero[3] = {0, 0, 0};
        pk::wb_backprojector(psf_fp, dims, bp_type, alpha, beta, n, res_flag, ires ? ires : zero, psf_bp, otf_bp, info);
    });
}

}  // extern "C"
