"""Minimal reader of the uncompressed little-endian float32 TIFF and BigTIFF
files the library writes (one strip per row or one strip per page)."""
import struct

import numpy as np


def read_tiff_f32(path, with_format=False):
    """Pages of the file -> float32 (pages, h, w); with_format: also 'tiff'
    or 'bigtiff'."""
    b = open(path, "rb").read()
    assert b[:2] == b"II"
    magic = struct.unpack_from("<H", b, 2)[0]
    assert magic in (42, 43)
    big = magic == 43
    if big:
        assert struct.unpack_from("<HH", b, 4) == (8, 0)
        off = struct.unpack_from("<Q", b, 8)[0]
    else:
        off = struct.unpack_from("<I", b, 4)[0]
    size = {3: 2, 4: 4, 5: 8, 16: 8}
    pages = []
    while off:
        if big:
            n = struct.unpack_from("<Q", b, off)[0]
            base, esz, fsz = off + 8, 20, 8
        else:
            n = struct.unpack_from("<H", b, off)[0]
            base, esz, fsz = off + 2, 12, 4
        tags = {}
        for i in range(n):
            e = base + esz * i
            tag, typ = struct.unpack_from("<HH", b, e)
            cnt = struct.unpack_from("<Q" if big else "<I", b, e + 4)[0]
            vpos = e + 4 + fsz
            nbytes = size[typ] * cnt
            if nbytes > fsz:  # out of place: the field holds the offset
                vpos = struct.unpack_from("<Q" if big else "<I", b, vpos)[0]
            fmt = {3: "H", 4: "I", 5: "II", 16: "Q"}[typ]
            vals = struct.unpack_from("<" + fmt * cnt, b, vpos)
            tags[tag] = (typ, cnt, vals)
        w, h = tags[256][2][0], tags[257][2][0]
        assert tags[258][2][0] == 32 and tags[339][2][0] == 3
        offs = tags[273][2]
        cnts = tags[279][2]
        rows = [np.frombuffer(b, "<f4", c // 4, o) for o, c in zip(offs, cnts)]
        pages.append(np.concatenate(rows).reshape(h, w))
        off = struct.unpack_from("<Q" if big else "<I", b, base + esz * n)[0]
    out = np.stack(pages)
    return (out, "bigtiff" if big else "tiff") if with_format else out
