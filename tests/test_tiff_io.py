"""TIFF stack I/O (readtiff / writetiff) and the LZW strip codec behind it.

The codec is pinned by the TIFF 6.0 specification's own LZW example (section
13: the bytes 7 7 7 8 8 7 7 6 6 encode to the codes 256 7 258 8 8 258 6 6 257),
by libtiff's LZW codec as Pillow exposes it (files written by one are read by
the other, both ways, at sizes that cross every code width and the table reset),
and by round trips.  CPU only: the codec is host code."""
import os
import struct

import numpy as np
import pytest

import petakit5d_amd.tiff_io as T

SPEC_IN = bytes([7, 7, 7, 8, 8, 7, 7, 6, 6])
SPEC_CODES = [256, 7, 258, 8, 8, 258, 6, 6, 257]


def _pack9(codes):
    bits = "".join(f"{c:09b}" for c in codes)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


def test_lzw_spec_example():
    assert T.lzw_encode(SPEC_IN) == _pack9(SPEC_CODES)
    assert T.lzw_decode(_pack9(SPEC_CODES), len(SPEC_IN)) == SPEC_IN


@pytest.mark.parametrize("kind", ["random", "runs", "image", "empty", "one"])
def test_lzw_round_trip(kind):
    g = np.random.default_rng(5)
    data = {
        "random": g.integers(0, 256, size=400_000, dtype=np.uint8).tobytes(),  # 12-bit codes and table resets
        "runs": bytes(np.repeat(g.integers(0, 4, size=5000, dtype=np.uint8), g.integers(1, 300, size=5000))),
        "image": (g.integers(0, 3000, size=300_000) // 5 * 5).astype(np.uint16).tobytes(),
        "empty": b"",
        "one": b"\x2a",
    }[kind]
    enc = T.lzw_encode(data)
    assert T.lzw_decode(enc, len(data)) == data


def test_lzw_errors():
    with pytest.raises(ValueError):
        T.lzw_decode(_pack9([256, 7, 400, 257]), 16)  # code beyond the table
    with pytest.raises(ValueError):
        T.lzw_decode(_pack9(SPEC_CODES), 4)  # output longer than the strip


@pytest.mark.parametrize("dtype", [np.uint16, np.float32, np.uint8])
@pytest.mark.parametrize("comp", ["lzw", "none"])
def test_write_read_round_trip(tmp_path, dtype, comp):
    g = np.random.default_rng(6)
    vol = (g.random((700, 45, 4)) * 3000).astype(dtype)  # 700 rows: two strips of 512 / 188 rows
    fn = str(tmp_path / "v.tif")
    T.writetiff(vol, fn, Compression=comp)
    r = T.readtiff(fn)
    assert r.dtype == vol.dtype and r.shape == vol.shape and r.flags.f_contiguous
    np.testing.assert_array_equal(r, vol)
    np.testing.assert_array_equal(T.readtiff(fn, range=[2, 3]), vol[:, :, 1:3])
    with pytest.raises(ValueError):
        T.readtiff(fn, range=[3, 5])


def test_two_dimensional(tmp_path):
    img = np.arange(35, dtype=np.uint16).reshape(5, 7)
    fn = str(tmp_path / "p.tif")
    T.writetiff(img, fn)
    np.testing.assert_array_equal(T.readtiff(fn)[:, :, 0], img)


def test_libtiff_interop(tmp_path):
    """Files written by libtiff (through Pillow) read here, and files written
    here read through libtiff, LZW and 16-bit, several pages."""
    Image = pytest.importorskip("PIL.Image")
    g = np.random.default_rng(7)
    pages = [(g.integers(0, 4000, size=(600, 257)) // 3 * 3).astype(np.uint16) for _ in range(3)]
    fn = str(tmp_path / "lib.tif")
    ims = [Image.fromarray(p) for p in pages]
    ims[0].save(fn, compression="tiff_lzw", save_all=True, append_images=ims[1:])
    r = T.readtiff(fn)
    np.testing.assert_array_equal(r, np.stack(pages, axis=2))
    fn2 = str(tmp_path / "mine.tif")
    T.writetiff(np.stack(pages, axis=2), fn2)
    with Image.open(fn2) as im:
        assert im.info.get("compression") == "tiff_lzw"
        for z, p in enumerate(pages):
            im.seek(z)
            np.testing.assert_array_equal(np.array(im), p)


def _classic_tiff(path, pages_bytes, W, H, bits, bo="<", predictor=None, desc=None, fmt=1):
    """A hand-assembled classic TIFF (one uncompressed strip per page)."""
    out = bytearray(b"II" if bo == "<" else b"MM")
    out += struct.pack(bo + "HI", 42, 0)
    link = 4
    for k, data in enumerate(pages_bytes):
        off = len(out)
        out += data
        if len(out) % 2:
            out += b"\0"
        ent = [(256, 4, W), (257, 4, H), (258, 3, bits), (259, 3, 1), (262, 3, 1), (273, 4, off),
               (277, 3, 1), (278, 4, H), (279, 4, len(data)), (339, 3, fmt)]
        if predictor:
            ent.append((317, 3, predictor))
        extra = b""
        if desc and k == 0:
            d = desc.encode() + b"\0"
            ent.append((270, 2, d))
        ent.sort(key=lambda e: e[0])
        ifd = len(out)
        xpos = ifd + 2 + 12 * len(ent) + 4
        body = b""
        for tag, typ, v in ent:
            if typ == 2:
                body += struct.pack(bo + "HHII", tag, typ, len(v), xpos + len(extra))
                extra += v
            elif typ == 3:
                body += struct.pack(bo + "HHIH", tag, typ, 1, v) + b"\0\0"
            else:
                body += struct.pack(bo + "HHII", tag, typ, 1, v)
        out += struct.pack(bo + "H", len(ent)) + body + b"\0\0\0\0" + extra
        struct.pack_into(bo + "I", out, link, ifd)
        link = ifd + 2 + 12 * len(ent)
    with open(path, "wb") as f:
        f.write(bytes(out))


def test_big_endian_and_predictor(tmp_path):
    img = np.array([[1, 5, 9, 2], [7, 7, 8, 0], [65535, 1, 3, 3]], dtype=np.uint16)
    diff = img.copy()
    diff[:, 1:] = (img[:, 1:].astype(np.int64) - img[:, :-1]) % 65536  # horizontal differencing
    fn = str(tmp_path / "be.tif")
    _classic_tiff(fn, [diff.astype(">u2").tobytes()], 4, 3, 16, bo=">", predictor=2)
    np.testing.assert_array_equal(T.readtiff(fn)[:, :, 0], img)


def test_imagej_stack(tmp_path):
    """ImageJ stacks keep one directory; the other pages follow its strip."""
    vol = np.arange(6 * 5 * 4, dtype=np.float32).reshape(4, 6, 5)  # pages, rows, columns
    fn = str(tmp_path / "ij.tif")
    _classic_tiff(fn, [vol.tobytes()], 5, 6, 32, desc="ImageJ=1.53t\nimages=4\nslices=4\n", fmt=3)
    r = T.readtiff(fn)
    assert r.shape == (6, 5, 4)
    np.testing.assert_array_equal(r, vol.transpose(1, 2, 0))
    np.testing.assert_array_equal(T.readtiff(fn, range=[3, 4]), vol[2:4].transpose(1, 2, 0))


def test_not_a_tiff(tmp_path):
    fn = str(tmp_path / "x.tif")
    with open(fn, "wb") as f:
        f.write(b"hello world")
    with pytest.raises(ValueError):
        T.readtiff(fn)
    assert not os.path.exists(str(tmp_path / "missing.tif"))
