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


class _Bits:
    def __init__(self, data: bytes):
        self.d, self.p = data, 0

    def get(self, n: int) -> int:
        v = 0
        for k in range(n):
            v |= ((self.d[self.p >> 3] >> (self.p & 7)) & 1) << k
            self.p += 1
        return v


def _decoder(lengths):
    """Canonical Huffman decoding table {(length, code): symbol} (RFC 1951 3.2.2)."""
    counts = [0] * 16
    for L in lengths:
        if L:
            counts[L] += 1
    code, first = 0, [0] * 16
    for L in range(1, 16):
        code = (code + counts[L - 1]) << 1 if L > 1 else 0
        first[L] = code
    table, nxt = {}, list(first)
    for s, L in enumerate(lengths):
        if L:
            table[(L, nxt[L])] = s
            nxt[L] += 1
    return table


def _read_sym(bits: _Bits, table) -> int:
    code, L = 0, 0
    while True:
        code = (code << 1) | bits.get(1)
        L += 1
        if (L, code) in table:
            return table[(L, code)]
        if L > 15:
            raise PngError("bad Huffman code")


def deflate_block_stats(png: bytes):
    """For a PNG whose IDAT holds a zlib stream of dynamic-Huffman blocks:
    (literal/length code lengths, distance code lengths, counts of every
    literal/length symbol) of the first block. A small inflater for tests."""
    idat = b""
    pos = 8
    while pos < len(png):
        n = struct.unpack(">I", png[pos:pos + 4])[0]
        if png[pos + 4:pos + 8] == b"IDAT":
            idat += png[pos + 8:pos + 8 + n]
        pos += 12 + n
    bits = _Bits(idat[2:])   # after the zlib header
    bits.get(1)
    if bits.get(2) != 2:
        raise PngError("not a dynamic block")
    hlit, hdist, hclen = bits.get(5) + 257, bits.get(5) + 1, bits.get(4) + 4
    order = [16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15]
    cl = [0] * 19
    for k in range(hclen):
        cl[order[k]] = bits.get(3)
    cl_table = _decoder(cl)
    lens = []
    while len(lens) < hlit + hdist:
        s = _read_sym(bits, cl_table)
        if s < 16:
            lens.append(s)
        elif s == 16:
            lens += [lens[-1]] * (3 + bits.get(2))
        elif s == 17:
            lens += [0] * (3 + bits.get(3))
        else:
            lens += [0] * (11 + bits.get(7))
    ll, dl = lens[:hlit], lens[hlit:]
    ll_table, d_table = _decoder(ll), _decoder(dl)
    counts = [0] * hlit
    extra = [0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0]
    dextra = [0, 0, 0, 0] + [k // 2 for k in range(2, 28)]
    while True:
        s = _read_sym(bits, ll_table)
        counts[s] += 1
        if s == 256:
            break
        if s > 256:
            bits.get(extra[s - 257])
            d = _read_sym(bits, d_table)
            bits.get(dextra[d])
    return ll, dl, counts
