"""The in-memory large-volume driver (RLdecon_large_in_memory.m:171-262 with
XR_zarrChunkCoordinatesExtraction) through one batched engine call: every
region of the output equals the single-volume call on its batch (border
round((psf + 10) / 2)) cropped to the region, bitwise, for host and device
volumes."""
import numpy as np
import pytest

from petakit5d_amd import synth

pytestmark = pytest.mark.gpu


def _vol(shape, seed):
    g = synth.rng(seed, 0)
    return np.asfortranarray(g.integers(80, 600, size=shape).astype(np.uint16))


@pytest.mark.parametrize("method,numit", [("omw", 2), ("rl", 3)])
def test_large_in_memory_equals_per_batch_calls(oracle, method, numit):
    import petakit5d_amd as pk
    from petakit5d_amd.chunks import XR_zarrChunkCoordinatesExtraction, matlab_round
    shape, bs = (150, 130, 60), (64, 64, 32)
    print(f"[large-in-memory] {method} {shape} batch {bs}", flush=True)
    im = _vol(shape, 5)
    psf = synth.gaussian_psf((9, 9, 11), 1.2, 3.0, 10.0)
    psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    J = pk.decon_large_in_memory(im, psf, psf_b, numit, method=method, batchSize=bs, EdgeErosion=0, Background=100,
                                 save16bit=False)
    assert J.shape == shape and J.dtype == np.float32
    border = matlab_round((np.array(psf.shape) + 10) / 2)
    batch, region, local = XR_zarrChunkCoordinatesExtraction(shape, batchSize=bs, BlockSize=bs, sameBatchSize=False,
                                                             BorderSize=border)
    assert len(batch) == 18
    for b, r, lo in zip(batch, region, local):
        c = np.asfortranarray(im[b[0] - 1:b[3], b[1] - 1:b[4], b[2] - 1:b[5]])
        pf, pb = pk.crop_psf_to_data(psf, c.shape), pk.crop_psf_to_data(psf_b, c.shape)
        if method == "omw":
            R, _ = pk.decon_lucy_omw_function(c, pf, pb, numit, Background=100, save16bit=False, pad=True)
        else:
            R, _, _ = pk.decon_lucy_function(c, pf, numit, Background=100, save16bit=False, pad=True)
        want = R[lo[0] - 1:lo[3], lo[1] - 1:lo[4], lo[2] - 1:lo[5]]
        np.testing.assert_array_equal(J[r[0] - 1:r[3], r[1] - 1:r[4], r[2] - 1:r[5]], want)
    # the first and the last batch against the oracle (batch zero-padded to
    # smooth sizes, result cropped back, region taken)
    from conftest import rel_l2
    for b, r, lo in (zip(batch, region, local).__next__(), (batch[-1], region[-1], local[-1])):
        c = np.asfortranarray(im[b[0] - 1:b[3], b[1] - 1:b[4], b[2] - 1:b[5]])
        pd = tuple(pk.findGoodFactorNumber(n, 1, False, maxPrime=5) for n in c.shape)
        P = np.zeros(pd, c.dtype, order="F")
        P[:c.shape[0], :c.shape[1], :c.shape[2]] = c
        pf, pb = pk.crop_psf_to_data(psf, c.shape), pk.crop_psf_to_data(psf_b, c.shape)
        if method == "omw":
            R, _ = oracle.decon_lucy_omw_function(P, pf, pb, numit, Background=100, save16bit=False)
        else:
            R, _, _ = oracle.decon_lucy_function(P, pf, numit, Background=100, save16bit=False)
        want = R[lo[0] - 1:lo[3], lo[1] - 1:lo[4], lo[2] - 1:lo[5]]
        e = rel_l2(J[r[0] - 1:r[3], r[1] - 1:r[4], r[2] - 1:r[5]], want)
        print(f"  batch {b}: region rel-L2 vs oracle FP32 {e:.3e}", flush=True)
        assert e <= 1e-4


def test_large_in_memory_device_volume_matches_host():
    import torch
    import petakit5d_amd as pk
    im = _vol((96, 80, 40), 6)
    psf = synth.gaussian_psf((7, 7, 9), 1.0, 2.5)
    psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    Jh = pk.decon_large_in_memory(im, psf, psf_b, 2, batchSize=(48, 48, 24), Background=100)
    dev = torch.from_numpy(np.ascontiguousarray(im.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    Jd = pk.decon_large_in_memory(dev, psf, psf_b, 2, batchSize=(48, 48, 24), Background=100)
    assert Jh.dtype == np.uint16
    np.testing.assert_array_equal(Jd.cpu().numpy(), Jh)


def _zarr_reference(im, psf, psf_b, numit, bs, ochunks):
    """The Zarr driver's result assembled from one decon call per batch on the
    driver's boxes (same-size batches, sorted by size, RLdecon_large_in_place.m:
    172-182)."""
    import petakit5d_amd as pk
    from petakit5d_amd.chunks import matlab_round
    from petakit5d_amd.zarr_io import zarr_batches
    bs = np.array([-(-int(x) // c) * c for x, c in zip(bs, ochunks)])
    border = matlab_round((np.array(psf.shape) + 10) / 2)
    batch, region, _ = zarr_batches(im.shape, bs, border)
    out = np.zeros(im.shape, np.uint16, order="F")
    for bb, r in zip(batch, region):
        c = np.asfortranarray(im[bb[0] - 1:bb[3], bb[1] - 1:bb[4], bb[2] - 1:bb[5]])
        R, _ = pk.decon_lucy_omw_function(c, pk.crop_psf_to_data(psf, c.shape), pk.crop_psf_to_data(psf_b, c.shape),
                                          numit, Background=100, pad=True)
        lo = r[:3] - bb[:3]
        hi = lo + (r[3:] - r[:3]) + 1
        out[r[0] - 1:r[3], r[1] - 1:r[4], r[2] - 1:r[5]] = R[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    return out, len(batch)


@pytest.mark.parametrize("compressor", ["zlib", "blosc"])
def test_decon_zarr_equals_per_batch_calls(tmp_path, compressor):
    """A Zarr store deconvolved group by group (same-size batches sorted by
    size, reads of the next group overlapping the device work, regions written
    straight into the output store) equals one decon call per batch on the same
    boxes, bitwise, whatever the group size; blosc stores (the reference's
    default codec) are written back as blosc."""
    import json
    import petakit5d_amd as pk
    print(f"[decon_zarr] {compressor}", flush=True)
    im = _vol((80, 72, 36), 8)
    psf = synth.gaussian_psf((7, 7, 9), 1.0, 2.5)
    psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    src = str(tmp_path / "in.zarr")
    pk.write_zarr(src, im, chunks=(32, 32, 16), compressor=compressor)
    R, _ = _zarr_reference(im, psf, psf_b, 2, (40, 40, 20), (32, 32, 16))
    for g in (16, 1, 3):
        dst = str(tmp_path / f"out{g}.zarr")
        # batchSize (40, 40, 20) rounds up to whole 32 x 32 x 16 chunks: (64, 64, 32)
        assert pk.decon_zarr(src, dst, psf, psf_b, 2, batchSize=(40, 40, 20), groupSize=g, Background=100) is None
        np.testing.assert_array_equal(pk.read_zarr(dst), R)
        assert json.load(open(f"{dst}/.zarray"))["compressor"]["id"] == compressor


def test_decon_zarr_flag_files_resume(tmp_path):
    """Groups that finished leave blocks_<first>_<last>.mat flag files
    (RLdecon_for_zarr_block.m:173-177); a second run skips them (:79-86), so
    a region erased from the output stays erased until Overwrite reruns it."""
    import os
    import glob
    import petakit5d_amd as pk
    im = _vol((80, 72, 36), 9)
    psf = synth.gaussian_psf((7, 7, 9), 1.0, 2.5)
    psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    src, dst = str(tmp_path / "in.zarr"), str(tmp_path / "out.zarr")
    pk.write_zarr(src, im, chunks=(32, 32, 16))
    R, nb = _zarr_reference(im, psf, psf_b, 2, (40, 40, 20), (32, 32, 16))
    pk.decon_zarr(src, dst, psf, psf_b, 2, batchSize=(40, 40, 20), groupSize=3, Background=100)
    flags = sorted(os.path.basename(f) for f in glob.glob(dst + ".zarr_flag/*.mat"))
    want = sorted(f"blocks_{a}_{min(a + 2, nb)}.mat" for a in range(1, nb + 1, 3))
    assert flags == want
    np.testing.assert_array_equal(pk.read_zarr(dst), R)
    victim = sorted(glob.glob(dst + "/0.0.0"))[0]
    os.remove(victim)  # an absent chunk reads as the fill value 0
    pk.decon_zarr(src, dst, psf, psf_b, 2, batchSize=(40, 40, 20), groupSize=3, Background=100)
    assert not pk.read_zarr(dst)[:32, :32, :16].any()
    pk.decon_zarr(src, dst, psf, psf_b, 2, batchSize=(40, 40, 20), groupSize=3, Background=100, Overwrite=True)
    np.testing.assert_array_equal(pk.read_zarr(dst), R)


@pytest.mark.parametrize("E", [8, 3])
def test_large_in_memory_edge_erosion(oracle, E, tmp_path):
    """The driver's default EdgeErosion = 8 (RLdecon_large_in_memory.m:25,
    141-162, 271-274): the assembled result is zeroed where bwdist(~bw) <=
    E - 1, bw = im > 0 with cleared faces; the mask is checked against
    scipy's exact Euclidean distance transform, the volume against the same
    run without erosion.  An ErodeMaskfile replaces the computed mask."""
    import torch
    from scipy import ndimage as ndi
    import petakit5d_amd as pk
    from petakit5d_amd.chunks import edge_erosion_mask
    im = _vol((96, 80, 40), 12)
    im[:30, :25, :] = 0   # an irregular I > 0 region
    im[60:64, 40:44, 10:30] = 0
    psf = synth.gaussian_psf((7, 7, 9), 1.0, 2.5)
    psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    bw = im > 0
    bw[[0, -1], :, :] = False
    bw[:, [0, -1], :] = False
    bw[:, :, [0, -1]] = False
    want = ndi.distance_transform_edt(bw) > E - 1
    m = edge_erosion_mask(im, E)
    np.testing.assert_array_equal(m, want)
    dev = torch.from_numpy(np.ascontiguousarray(im.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    np.testing.assert_array_equal(edge_erosion_mask(dev, E).cpu().numpy(), want)
    J0 = pk.decon_large_in_memory(im, psf, psf_b, 2, batchSize=(48, 48, 24), EdgeErosion=0, Background=100)
    JE = pk.decon_large_in_memory(im, psf, psf_b, 2, batchSize=(48, 48, 24), EdgeErosion=E, Background=100)
    np.testing.assert_array_equal(JE, np.where(want, J0, 0))
    fn = str(tmp_path / "mask.tif")
    pk.writetiff(want[:, ::-1, :].astype(np.uint8).copy(order="F"), fn)
    JM = pk.decon_large_in_memory(im, psf, psf_b, 2, batchSize=(48, 48, 24), ErodeMaskfile=fn, Background=100)
    np.testing.assert_array_equal(JM, np.where(want[:, ::-1, :], J0, 0))
