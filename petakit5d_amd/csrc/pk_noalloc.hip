// pk_noalloc.hip -- decon_lucy_omw_function for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  Every device buffer of the call -- plan tables,
// OTF_f and OTF_b, I, the estimate, the spectrum, the erosion masks -- is a
// piece of one static device array per GPU, `scratch`.
//
// `scratch` keeps the last call's contents and every call on a GPU shares
// it: calls that overlap on one GPU (two threads, two contexts on the same
// device) race.  Nothing read from it is assumed to start at zero; the call
// fills what it reads before reading it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/pkdecon.h"
#include "pk_engine.hpp"

struct pk_ctx : pk::pk_ctx_impl {};

// In host code this name is a host-side stand-in, not a device address: the
// device address of each GPU's copy comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[PK_NOALLOC_SCRATCH_BYTES];

namespace {
struct DevScratch {
    void* ptr = nullptr;  // this device's copy of `scratch`
    int state = 0;        // 0: not resolved, 1: resolved, -1: hipGetSymbolAddress failed
    hipError_t err = hipSuccess;
};
std::mutex g_mu;
std::vector<DevScratch> g_dev;

// Resolves `scratch` on `device`, which must be the current device; once per
// device, later calls return the kept result.
DevScratch resolve(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_dev.size() <= device) g_dev.resize((size_t)device + 1);
    DevScratch& d = g_dev[(size_t)device];
    if (d.state == 0) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            d.state = -1;
            d.err = e != hipSuccess ? e : hipErrorInvalidSymbol;
        } else {
            d.ptr = p;
            d.state = 1;
        }
    }
    return d;
}
}  // namespace

namespace pk {
// Set-up of the scratch array on a device (pk_ctx_create, with the device current).
void noalloc_setup(int device) { (void)resolve(device); }
}  // namespace pk

extern "C" pk_status pk_decon_lucy_omw_noalloc(pk_ctx* c, const pk_vol* I, const float* psf_f,
                                               const int64_t f_dims[3], const float* psf_b, const int64_t b_dims[3],
                                               int32_t numit, const pk_opts* opts, pk_vol* out, double* err_mat) {
    if (!c) return PK_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    try {
        PK_HIP(hipSetDevice(c->device));
        if (!f_dims || !b_dims) pk::fail(PK_ERR_INVALID_ARG, "null PSF size");
        const DevScratch d = resolve(c->device);
        if (d.state != 1)
            pk::fail(PK_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(d.err));
        c->decon_omw_noalloc(I, psf_f, f_dims, psf_b, b_dims, numit, opts, out, err_mat, d.ptr,
                             (size_t)PK_NOALLOC_SCRATCH_BYTES);
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
