"""TIFF stack I/O with the reference's conventions: readtiff / writetiff
(microscopeDataProcessing/io/readtiff.m, writetiff.m over io/cpp-tiff).

A stack is one image per directory (page); the array comes back in MATLAB's
size order [y x z] (rows, columns, pages), column-major.  Strips are
uncompressed or LZW, the LZW strips decoded and encoded by libpkdecon.so's
host codec (pk_lzw_decode / pk_lzw_encode) on a pool of host threads, as the
reference spreads strips over OpenMP workers.  The writer follows
parallelwritetiff.cpp:38-60: 512 rows per strip (:321), LZW unless
Compression='none', PHOTOMETRIC 1, PLANARCONFIG 1, one sample per pixel,
SAMPLEFORMAT IEEEFP for 32/64-bit data; BigTIFF when the file passes 4 GiB.
The reader also takes classic or BigTIFF in either byte order, predictor 2
(horizontal differencing) and ImageJ stacks whose pages follow the first one
without directories of their own (helperfunctions.cpp:76-120)."""
from __future__ import annotations

import ctypes
import os
import re
import struct
from concurrent.futures import ThreadPoolExecutor

import numpy as np

__all__ = ["readtiff", "writetiff", "lzw_decode", "lzw_encode"]

_ROWS_PER_STRIP = 512
_SAMPLE_FORMAT = {1: "u", 2: "i", 3: "f"}


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _codec():
    from ._lib import load
    L = load()
    for name in ("pk_lzw_decode", "pk_lzw_encode"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                      ctypes.POINTER(ctypes.c_size_t)]
    L.pk_lzw_encode_bound.restype = ctypes.c_size_t
    L.pk_lzw_encode_bound.argtypes = [ctypes.c_size_t]
    return L


def lzw_decode(data: bytes, size: int) -> bytes:
    """One TIFF LZW strip -> at most size bytes (libpkdecon.so's pk_lzw_decode)."""
    L = _codec()
    src = np.frombuffer(data, dtype=np.uint8)
    dst = np.empty(max(size, 1), dtype=np.uint8)
    n = ctypes.c_size_t(0)
    rc = L.pk_lzw_decode(src.ctypes.data if src.size else None, src.size, dst.ctypes.data, size, ctypes.byref(n))
    if rc != 0:
        raise ValueError(f"LZW strip: {'output larger than the strip' if rc == 2 else 'malformed code stream'}")
    return dst[:n.value].tobytes()


def lzw_encode(data) -> bytes:
    """Bytes -> one TIFF LZW strip (libpkdecon.so's pk_lzw_encode)."""
    L = _codec()
    src = np.frombuffer(memoryview(data).cast("B"), dtype=np.uint8)
    cap = int(L.pk_lzw_encode_bound(src.size))
    dst = np.empty(cap, dtype=np.uint8)
    n = ctypes.c_size_t(0)
    rc = L.pk_lzw_encode(src.ctypes.data if src.size else None, src.size, dst.ctypes.data, cap, ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f"pk_lzw_encode failed ({rc})")
    return dst[:n.value].tobytes()


# ---------------------------------------------------------------- reading

_TYPES = {1: "B", 2: "s", 3: "H", 4: "I", 5: "II", 6: "b", 7: "B", 8: "h", 9: "i", 10: "ii", 11: "f", 12: "d",
          16: "Q", 17: "q", 18: "Q"}


def _ifds(buf):
    bo = {b"II": "<", b"MM": ">"}.get(bytes(buf[:2]))
    if bo is None:
        raise ValueError("not a TIFF file")
    ver = struct.unpack_from(bo + "H", buf, 2)[0]
    if ver == 42:
        big, off = False, struct.unpack_from(bo + "I", buf, 4)[0]
    elif ver == 43:
        big, off = True, struct.unpack_from(bo + "Q", buf, 8)[0]
    else:
        raise ValueError("not a TIFF file")
    pages = []
    seen = set()
    while off and off not in seen:
        seen.add(off)
        n = struct.unpack_from(bo + ("Q" if big else "H"), buf, off)[0]
        ent = off + (8 if big else 2)
        esz = 20 if big else 12
        tags = {}
        for i in range(n):
            e = ent + i * esz
            if big:
                tag, typ, cnt = struct.unpack_from(bo + "HHQ", buf, e)
                vo, slot = e + 12, 8
            else:
                tag, typ, cnt = struct.unpack_from(bo + "HHI", buf, e)
                vo, slot = e + 8, 4
            code = _TYPES.get(typ)
            if code is None:
                continue
            if code == "s":
                size = cnt
            else:
                size = struct.calcsize(bo + code) * cnt
            if size > slot:
                vo = struct.unpack_from(bo + ("Q" if big else "I"), buf, vo)[0]
            if code == "s":
                tags[tag] = bytes(buf[vo:vo + cnt]).rstrip(b"\0").decode("latin-1")
            else:
                v = struct.unpack_from(bo + code * cnt, buf, vo)
                tags[tag] = list(v)
        pages.append(tags)
        off = struct.unpack_from(bo + ("Q" if big else "I"), buf, ent + n * esz)[0]
    return bo, pages


def _page_layout(t, bo):
    W, H = int(t[256][0]), int(t[257][0])
    bits = int(t.get(258, [1])[0])
    if int(t.get(277, [1])[0]) != 1:
        raise ValueError("only one sample per pixel is supported")
    fmt = _SAMPLE_FORMAT.get(int(t.get(339, [1])[0]))
    if fmt is None or bits not in (8, 16, 32, 64):
        raise ValueError(f"unsupported sample format / {bits} bits")
    return W, H, np.dtype(f"{bo}{fmt}{bits // 8}")


def _decode_page(buf, t, bo):
    W, H, dt = _page_layout(t, bo)
    comp = int(t.get(259, [1])[0])
    if comp not in (1, 5):
        raise ValueError(f"TIFF compression {comp} is not supported (none and LZW are)")
    offs, cnts = t[273], t.get(279)
    rps = int(t.get(278, [H])[0])
    row_bytes = W * dt.itemsize
    parts = []
    for i, o in enumerate(offs):
        rows = min(rps, H - i * rps)
        want = rows * row_bytes
        raw = bytes(buf[o:o + (cnts[i] if cnts else want)])
        if comp == 5:
            raw = lzw_decode(raw, want)
        if len(raw) < want:
            raise ValueError("TIFF strip shorter than its rows")
        parts.append(raw[:want])
    img = np.frombuffer(b"".join(parts), dtype=dt).reshape(H, W)
    if int(t.get(317, [1])[0]) == 2:  # horizontal differencing
        if dt.kind not in "ui":
            raise ValueError("predictor 2 on floating-point samples is not supported")
        img = np.cumsum(img, axis=1, dtype=dt)
    return img.astype(dt.newbyteorder("="), copy=False)


def readtiff(filepath, range=None):
    """data = readtiff(filepath, 'range', [zStart zEnd]) (readtiff.m:1-30):
    pages zStart..zEnd (1-based, inclusive; all pages when None) as a
    column-major [y x z] array of the file's sample type."""
    with open(filepath, "rb") as f:
        buf = memoryview(f.read())
    bo, pages = _ifds(buf)
    if not pages:
        raise ValueError("TIFF file without images")
    desc = pages[0].get(270, "")
    nz_ij = None
    if isinstance(desc, str) and "ImageJ" in desc and len(pages) == 1 and int(pages[0].get(259, [1])[0]) == 1:
        m = re.search(r"images=(\d+)", desc)
        if m and int(m.group(1)) > 1:
            nz_ij = int(m.group(1))
    nz = nz_ij or len(pages)
    z0, z1 = (0, nz) if range is None else (int(range[0]) - 1, int(range[1]))
    if z0 < 0 or z1 > nz or z0 >= z1:
        raise ValueError(f"range {range} outside 1..{nz}")
    if nz_ij:  # ImageJ stack: pages stored back to back from the first strip on
        W, H, dt = _page_layout(pages[0], bo)
        start = int(pages[0][273][0]) + z0 * H * W * dt.itemsize
        cnt = (z1 - z0) * H * W
        st = np.frombuffer(buf, dtype=dt, count=cnt, offset=start).reshape(z1 - z0, H, W)
        st = st.astype(dt.newbyteorder("="), copy=False)
    else:
        with ThreadPoolExecutor(_threads()) as ex:
            imgs = list(ex.map(lambda t: _decode_page(buf, t, bo), pages[z0:z1]))
        if len({im.shape for im in imgs}) != 1 or len({im.dtype for im in imgs}) != 1:
            raise ValueError("pages differ in size or sample type")
        st = np.stack(imgs)
    return np.asfortranarray(st.transpose(1, 2, 0))


# ---------------------------------------------------------------- writing

def writetiff(img, filepath, Compression="lzw"):
    """writetiff(img, filepath, 'Compression', 'lzw' | 'none') (writetiff.m:1-60
    with parallelwritetiff.cpp:38-60): one page per z plane, 512 rows per strip."""
    if Compression.lower() not in ("lzw", "none"):
        raise ValueError("Compression must be 'lzw' or 'none'")
    a = np.asarray(img)
    if a.ndim == 2:
        a = a[:, :, None]
    if a.ndim != 3:
        raise ValueError("img must be 2-D or 3-D")
    if a.dtype.kind == "b":
        a = a.astype(np.uint8)
    if a.dtype.kind not in "uif" or a.dtype.itemsize not in (1, 2, 4, 8):
        raise ValueError(f"unsupported sample type {a.dtype}")
    a = a.astype(a.dtype.newbyteorder("<"), copy=False)
    H, W, Z = a.shape
    lzw = Compression.lower() == "lzw"
    rps = min(_ROWS_PER_STRIP, max(H, 1))
    nstrip = -(-H // rps) if H else 0

    def strips(z):
        page = np.ascontiguousarray(a[:, :, z])
        out = []
        for i in range(nstrip):
            raw = page[i * rps:(i + 1) * rps].tobytes()
            out.append(lzw_encode(raw) if lzw else raw)
        return out

    with ThreadPoolExecutor(_threads()) as ex:
        data = list(ex.map(strips, range(Z)))
    total = sum(len(s) for p in data for s in p)
    big = total + Z * (nstrip * 16 + 512) > 0xFFFF0000
    bo = "<"
    fmt_sample = {"u": 1, "i": 2, "f": 3}[a.dtype.kind]
    bits = a.dtype.itemsize * 8
    os.makedirs(os.path.dirname(os.path.abspath(filepath)), exist_ok=True)
    tmp = filepath + f".{os.getpid()}.tmp"
    with open(tmp, "wb") as f:
        if big:
            f.write(b"II" + struct.pack(bo + "HHHQ", 43, 8, 0, 0))
        else:
            f.write(b"II" + struct.pack(bo + "HI", 42, 0))
        prev_link = 8 if big else 4  # where the next IFD offset goes
        for z in range(Z):
            offs, cnts = [], []
            for s in data[z]:
                offs.append(f.tell())
                cnts.append(len(s))
                f.write(s)
            if f.tell() % 2:
                f.write(b"\0")
            ent = [(256, 4, [W]), (257, 4, [H]), (258, 3, [bits]), (259, 3, [5 if lzw else 1]), (262, 3, [1]),
                   (273, 16 if big else 4, offs), (277, 3, [1]), (278, 4, [rps]),
                   (279, 16 if big else 4, cnts), (284, 3, [1])]
            if bits >= 32 or fmt_sample != 1:
                ent.append((339, 3, [fmt_sample]))
            ifd_pos = f.tell()
            esz, slot = (20, 8) if big else (12, 4)
            extra_pos = ifd_pos + (8 if big else 2) + len(ent) * esz + (8 if big else 4)
            head = struct.pack(bo + ("Q" if big else "H"), len(ent))
            body, extra = b"", b""
            for tag, typ, vals in ent:
                code = _TYPES[typ]
                payload = struct.pack(bo + code * len(vals), *vals)
                if len(payload) <= slot:
                    val = payload.ljust(slot, b"\0")
                else:
                    val = struct.pack(bo + ("Q" if big else "I"), extra_pos + len(extra))
                    extra += payload
                body += struct.pack(bo + ("HHQ" if big else "HHI"), tag, typ, len(vals)) + val
            f.write(head + body + b"\0" * (8 if big else 4) + extra)
            end = f.tell()
            f.seek(prev_link)
            f.write(struct.pack(bo + ("Q" if big else "I"), ifd_pos))
            f.seek(end)
            prev_link = ifd_pos + (8 if big else 2) + len(ent) * esz
    os.replace(tmp, filepath)
