"""A small PNG reader for the tests: chunk CRCs checked with zlib.crc32, the
zlib stream inflated by zlib (which checks Adler-32), rows unfiltered per the
PNG specification. Test infrastructure only."""
from __future__ import annotations

import struct
import zlib

import numpy as np


class PngError(ValueError):
    pass


def read_png(data: bytes):
    """(pixels (h, w, channels) uint8, info dict) of an 8-bit RGB/RGBA PNG."""
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise PngError("signature")
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        if zlib.crc32(typ + body) & 0xFFFFFFFF != crc:
            raise PngError(f"CRC of {typ!r}")
        chunks.append((typ, body))
        pos += 12 + n
        if typ == b"IEND":
            break
    if pos != len(data):
        raise PngError("bytes after IEND")
    if chunks[0][0] != b"IHDR" or chunks[-1][0] != b"IEND":
        raise PngError("chunk order")
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    if depth != 8 or ctype not in (2, 6) or comp or filt or inter:
        raise PngError("unsupported IHDR")
    ch = 3 if ctype == 2 else 4
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    stride = w * ch
    if len(raw) != h * (stride + 1):
        raise PngError("inflated size")
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    filters = []
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        if f in (0, 2):
            cur = (line + (prev if f == 2 else 0)) & 255
        else:
            for i in range(stride):
                a = cur[i - ch] if i >= ch else 0
                b = prev[i]
                c = prev[i - ch] if i >= ch else 0
                if f == 1:
                    pr = a
                elif f == 3:
                    pr = (a + b) >> 1
                elif f == 4:
                    p = a + b - c
                    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                    pr = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                else:
                    raise PngError(f"filter {f}")
                cur[i] = (line[i] + pr) & 255
        filters.append(f)
        out[y] = cur
        prev = cur
    return out.reshape(h, w, ch), {"width": w, "height": h, "channels": ch, "filters": filters,
                                   "chunks": [t for t, _ in chunks]}


def write_png(pixels: np.ndarray, filt: int = 0) -> bytes:
    """A plain PNG writer (one filter type for all rows, zlib level 6): the
    reader's own check."""
    h, w, ch = pixels.shape
    rows = []
    prev = np.zeros(w * ch, np.int32)
    for y in range(h):
        cur = pixels[y].reshape(-1).astype(np.int32)
        if filt == 0:
            line = cur
        elif filt == 2:
            line = (cur - prev) & 255
        elif filt == 1:
            left = np.concatenate([np.zeros(ch, np.int32), cur[:-ch]])
            line = (cur - left) & 255
        else:
            raise ValueError(filt)
        rows.append(bytes([filt]) + line.astype(np.uint8).tobytes())
        prev = cur

    def chunk(t, b):
        return struct.pack(">I", len(b)) + t + b + struct.pack(">I", zlib.crc32(t + b) & 0xFFFFFFFF)

    ihdr = struct.pack(">IIBBBBB", w, h, 8, 2 if ch == 3 else 6, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(b"".join(rows), 6)) +
            chunk(b"IEND", b""))
