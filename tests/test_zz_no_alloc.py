"""pk_decon_lucy_omw_noalloc (the OMW entry point that allocates nothing:
every device buffer is a piece of the library's static per-GPU scratch
array) against pk_decon_lucy_omw through the C-ABI: two calls in a row, the
second on a smaller volume with other options, each bitwise equal to the
allocating entry point's result; then the limits (host volumes, sizes beyond
the scratch array) are rejected with their errors."""
import ctypes

import numpy as np
import pytest

from petakit5d_amd import synth

pytestmark = pytest.mark.gpu


def _vol(shape, seed):
    psf = synth.gaussian_psf((11, 11, 13), 1.2, 3.0, 20.0)
    g0, g1 = synth.rng(seed, 0), synth.rng(seed, 1)
    img = synth._fft_convolve_circular(synth.point_emitters(shape, 60, (500, 3000), g0), psf) + 100.0
    I = synth.poisson_u16(img, g1)
    I[:2, :, :] = 0  # an irregular I ~= 0 mask for the offset / erosion branches
    return I, psf


def _call(L, ctx, fn, It, psf, pb, numit, o, out_shape, out_dtype):
    import torch
    from petakit5d_amd import _lib
    d = tuple(It.shape)
    out = torch.empty((out_shape[2], out_shape[1], out_shape[0]), dtype=out_dtype, device="cuda").permute(2, 1, 0)
    code_in = _lib.PK_UINT16 if It.dtype == torch.uint16 else _lib.PK_SINGLE
    vin = _lib.PkVol(ctypes.c_void_p(It.data_ptr()), code_in, _lib.PK_MEM_DEVICE, (ctypes.c_int64 * 3)(*d))
    vout = _lib.PkVol(ctypes.c_void_p(out.data_ptr()), _lib.PK_UINT16 if out_dtype == torch.uint16 else _lib.PK_SINGLE,
                      _lib.PK_MEM_DEVICE, (ctypes.c_int64 * 3)(*out_shape))
    pf = np.asfortranarray(psf, dtype=np.float32)
    pbb = np.asfortranarray(pb, dtype=np.float32)
    err = np.full((max(numit, 1), 4), 7.0, order="F")
    torch.cuda.synchronize()
    st = getattr(L, fn)(ctx.handle, ctypes.byref(vin), ctypes.c_void_p(pf.ctypes.data), _lib.i64x3(pf.shape),
                        ctypes.c_void_p(pbb.ctypes.data), _lib.i64x3(pbb.shape), numit, ctypes.byref(o),
                        ctypes.byref(vout), ctypes.c_void_p(err.ctypes.data))
    return st, out, err


def _opts(L, **kw):
    from petakit5d_amd import _lib
    o = _lib.PkOpts()
    L.pk_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        if k == "bbox":
            o.has_bbox = 1
            for i in range(6):
                o.bbox[i] = v[i]
        else:
            setattr(o, k, v)
    return o


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).cuda().permute(2, 1, 0)


def test_noalloc_matches_allocating_entry_point():
    import torch
    import petakit5d_amd as pk
    from petakit5d_amd import _lib
    L = _lib.load()
    ctx = pk.Context(0)
    cases = [  # (shape, numit, options, output class): the second call is smaller
        ((96, 80, 60), 2, dict(background=100.0, save16bit=0), torch.float32, None),
        ((64, 48, 30), 3, dict(background=100.0, save16bit=1, edge_erosion=2, decon_offset=1.0, scale_factor=2.0),
         torch.uint16, [3, 2, 2, 60, 45, 28]),
        ((45, 36, 20), 1, dict(background=99.5, save16bit=0, damp_factor=1.5), torch.float32, None),  # odd rows
    ]
    for k, (shape, numit, kw, odt, bbox) in enumerate(cases):
        print(f"[noalloc] call {k}: {shape} numit={numit} {kw} bbox={bbox}", flush=True)
        I, psf = _vol(shape, 90 + k)
        pb = (np.ascontiguousarray(psf[::-1, ::-1, ::-1])
              * (1 + 0.2 * np.random.default_rng(k).standard_normal(psf.shape))).astype(np.float32)
        It = _cuda(I) if k != 2 else _cuda(I.astype(np.float32) + 0.25)
        if bbox:
            kw = dict(kw, bbox=bbox)
        os_ = tuple(bbox[i + 3] - bbox[i] + 1 for i in range(3)) if bbox else shape
        st_a, A, err_a = _call(L, ctx, "pk_decon_lucy_omw_noalloc", It, psf, pb, numit, _opts(L, **kw), os_, odt)
        assert st_a == 0, L.pk_last_error(ctx.handle)
        st_b, B, _ = _call(L, ctx, "pk_decon_lucy_omw", It, psf, pb, numit, _opts(L, **kw), os_, odt)
        assert st_b == 0, L.pk_last_error(ctx.handle)
        torch.cuda.synchronize()
        assert torch.equal(A, B), f"call {k}: no-allocation result differs from pk_decon_lucy_omw"
        assert not err_a[:numit].any()
    ctx.close()


def test_noalloc_rejects_what_it_cannot_serve():
    import torch
    import petakit5d_amd as pk
    from petakit5d_amd import _lib
    L = _lib.load()
    ctx = pk.Context(0)
    psf = synth.gaussian_psf((7, 7, 9), 1.0, 2.5)
    # 320 x 320 x 160 needs ~330 MB, beyond the 128 MiB scratch array
    big = torch.zeros((160, 320, 320), dtype=torch.uint16, device="cuda").permute(2, 1, 0)
    st, _, _ = _call(L, ctx, "pk_decon_lucy_omw_noalloc", big, psf, psf, 2, _opts(L, background=0.0, save16bit=1),
                     (320, 320, 160), torch.uint16)
    assert st == 3 and b"device scratch" in L.pk_last_error(ctx.handle)
    # host volumes are not accepted
    I = np.zeros((32, 32, 16), np.uint16, order="F")
    vin = _lib.PkVol(ctypes.c_void_p(I.ctypes.data), _lib.PK_UINT16, _lib.PK_MEM_HOST, (ctypes.c_int64 * 3)(*I.shape))
    O = np.zeros((32, 32, 16), np.uint16, order="F")
    vout = _lib.PkVol(ctypes.c_void_p(O.ctypes.data), _lib.PK_UINT16, _lib.PK_MEM_HOST, (ctypes.c_int64 * 3)(*O.shape))
    pf = np.asfortranarray(psf)
    st = L.pk_decon_lucy_omw_noalloc(ctx.handle, ctypes.byref(vin), ctypes.c_void_p(pf.ctypes.data),
                                     _lib.i64x3(pf.shape), ctypes.c_void_p(pf.ctypes.data), _lib.i64x3(pf.shape), 2,
                                     ctypes.byref(_opts(L)), ctypes.byref(vout), None)
    assert st == 1 and b"device volumes only" in L.pk_last_error(ctx.handle)
    # after the rejections a call that fits still succeeds
    small = _cuda(_vol((32, 32, 16), 5)[0])
    st, _, _ = _call(L, ctx, "pk_decon_lucy_omw_noalloc", small, psf, psf, 2, _opts(L, background=100.0, save16bit=1),
                     (32, 32, 16), torch.uint16)
    assert st == 0
    ctx.close()
