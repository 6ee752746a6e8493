"""Chunk coordinates of the large-volume drivers
(stitch/XR_zarrChunkCoordinatesExtraction.m:1-100), pinned by hand-derived
cases and by invariants a dropped term or a swapped index would break: the
regions tile the processed world exactly once, every batch holds its region
plus the border (clipped to the volume), local boxes map each region into its
batch, and with sameBatchSize every batch spans batchSize before borders."""
import itertools

import numpy as np
import pytest

from petakit5d_amd.chunks import XR_zarrChunkCoordinatesExtraction as coords, matlab_round


def test_hand_derived_1d_like_case():
    # 10 x 1 x 1 volume, batches of 4 (blocks of 4), border 1
    b, r, l = coords((10, 1, 1), batchSize=(4, 4, 4), BlockSize=(4, 4, 4), sameBatchSize=False, BorderSize=1)
    np.testing.assert_array_equal(r, [[1, 1, 1, 4, 1, 1], [5, 1, 1, 8, 1, 1], [9, 1, 1, 10, 1, 1]])
    np.testing.assert_array_equal(b, [[1, 1, 1, 5, 1, 1], [4, 1, 1, 9, 1, 1], [8, 1, 1, 10, 1, 1]])
    np.testing.assert_array_equal(l, [[1, 1, 1, 4, 1, 1], [2, 1, 1, 5, 1, 1], [2, 1, 1, 3, 1, 1]])
    # sameBatchSize: the last batch starts batchSize - 1 before its region's end
    b, r, l = coords((10, 1, 1), batchSize=(4, 4, 4), BlockSize=(4, 4, 4), sameBatchSize=True, BorderSize=1)
    np.testing.assert_array_equal(b[2], [6, 1, 1, 10, 1, 1])
    np.testing.assert_array_equal(l[2], [4, 1, 1, 5, 1, 1])


def test_batch_rounded_to_whole_blocks_and_column_major_order():
    # batchSize 5 with blocks of 4 -> 8; 20 -> regions [1-8], [9-16], [17-20]; dim 1 fastest
    b, r, _ = coords((20, 12, 3), batchSize=(5, 5, 5), BlockSize=(4, 4, 4), sameBatchSize=False, BorderSize=0)
    assert r.shape == (3 * 2 * 1, 6)
    np.testing.assert_array_equal(r[:4, :3], [[1, 1, 1], [9, 1, 1], [17, 1, 1], [1, 9, 1]])
    np.testing.assert_array_equal(r[2, 3:], [20, 8, 3])


@pytest.mark.parametrize("vol,bs,blk,same,border,bbox", [
    ((37, 29, 11), (10, 8, 5), (5, 4, 5), False, (3, 2, 1), None),
    ((37, 29, 11), (10, 8, 5), (5, 4, 5), True, (3, 2, 1), None),
    ((64, 40, 25), (32, 32, 32), (16, 16, 16), True, 7, None),
    ((50, 44, 30), (16, 12, 9), (4, 6, 3), True, (5, 5, 5), (3, 5, 2, 47, 40, 29)),
    ((50, 44, 30), (100, 100, 100), (256, 256, 256), False, 4, None),
])
def test_invariants(vol, bs, blk, same, border, bbox):
    b, r, l = coords(vol, batchSize=bs, BlockSize=blk, bbox=bbox, sameBatchSize=same, BorderSize=border)
    vol = np.array(vol)
    ws = np.array(bbox[:3]) if bbox else np.ones(3, int)
    we = np.array(bbox[3:]) if bbox else vol
    wsize = we - ws + 1
    bdr = np.broadcast_to(np.asarray(border), (3,))
    cover = np.zeros(tuple(wsize), dtype=np.int32)
    bsz = np.minimum(wsize, bs)
    bsz = -(-bsz // np.minimum(wsize, blk)) * np.minimum(wsize, blk)
    for bi, ri, li in zip(b, r, l):
        ra = ri[:3] + ws - 1, ri[3:] + ws - 1  # absolute region
        cover[tuple(slice(s - 1, e) for s, e in zip(ri[:3], ri[3:]))] += 1
        # batch inside the volume, region (absolute) inside the batch with the border where the volume allows
        assert np.all(bi[:3] >= 1) and np.all(bi[3:] <= vol)
        assert np.all(bi[:3] <= ra[0]) and np.all(bi[3:] >= ra[1])
        assert np.all(bi[3:] == np.minimum(vol, ra[1] + bdr))
        if not same:
            assert np.all(bi[:3] == np.maximum(1, ra[0] - bdr))
        else:
            assert np.all(bi[:3] == np.maximum(1, ra[1] - bsz + 1 - bdr))
        # local box: the region's place inside the batch, same extent
        np.testing.assert_array_equal(li[:3], ra[0] - bi[:3] + 1)
        np.testing.assert_array_equal(li[3:] - li[:3], ri[3:] - ri[:3])
    assert cover.min() == 1 and cover.max() == 1  # exact tiling of the world


def test_matlab_round_half_away_from_zero():
    np.testing.assert_array_equal(matlab_round([0.5, 1.5, 2.5, -0.5, 37.5, 37.4]), [1, 2, 3, -1, 38, 37])
    # the driver's border for a 64x64x101 PSF: round((psf + 10) / 2)
    np.testing.assert_array_equal(matlab_round((np.array([64, 64, 101]) + 10) / 2), [37, 37, 56])


def test_invalid_volume_size():
    with pytest.raises(ValueError):
        coords((0, 10, 10))
