"""Host checks of the PNG test reader (so the GPU encoder's tests rest on a
pinned decoder), and of the reference's fpng output through it."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from tests.png_util import PngError, read_png, write_png

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libmkref.so"


def _img(h, w, ch, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (h, w, ch), dtype=np.uint8)
    a[: h // 2, : w // 2] = 0   # a flat corner
    return a


@pytest.mark.parametrize("filt", [0, 1, 2])
@pytest.mark.parametrize("shape", [(1, 1, 3), (5, 7, 4), (33, 17, 3), (64, 64, 4)])
def test_reader_round_trip(filt, shape):
    a = _img(*shape, seed=sum(shape) + filt)
    got, info = read_png(write_png(a, filt))
    assert np.array_equal(got, a)
    assert info["filters"] == [filt] * shape[0]


def test_reader_rejects_corruption():
    png = bytearray(write_png(_img(8, 8, 3, 1)))
    png[40] ^= 1   # inside IDAT
    with pytest.raises((PngError, Exception)):
        read_png(bytes(png))


@pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("shape", [(32, 48, 3), (17, 9, 4), (128, 128, 3)])
def test_reader_decodes_fpng(shape):
    lib = C.CDLL(str(REF))
    lib.ref_fpng_encode.restype = C.c_ulonglong
    lib.ref_fpng_encode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_ulonglong]
    a = np.ascontiguousarray(_img(*shape, seed=5))
    out = np.zeros(a.size * 2 + 1024, np.uint8)
    n = lib.ref_fpng_encode(a.ctypes.data, shape[1], shape[0], shape[2], out.ctypes.data, out.size)
    assert n > 0
    got, _ = read_png(out[:n].tobytes())
    assert np.array_equal(got, a)


def test_block_reader_counts_literals_of_a_huffman_only_stream():
    """The tests' small inflater (deflate_block_stats) against zlib: a
    Huffman-only stream (no matches) of one dynamic block holds exactly the
    data's bytes as literals, and its code is complete."""
    import struct
    import zlib

    from tests.png_util import deflate_block_stats

    rng = np.random.default_rng(4)
    data = (np.abs(rng.normal(0, 12, 6000)).astype(np.uint8)).tobytes()
    co = zlib.compressobj(9, zlib.DEFLATED, 15, 9, zlib.Z_HUFFMAN_ONLY)
    z = co.compress(data) + co.flush()
    png = b"\x89PNG\r\n\x1a\n" + struct.pack(">I", len(z)) + b"IDAT" + z + b"\0\0\0\0"
    ll, dl, counts = deflate_block_stats(png)
    hist = np.bincount(np.frombuffer(data, np.uint8), minlength=256)
    assert counts[:256] == hist.tolist() and counts[256] == 1 and sum(counts[257:]) == 0
    assert sum(2.0 ** -L for L in ll if L) == 1.0
