"""Device PNG + base64 encoding: every PNG decodes (chunk CRCs, Adler-32,
filters) to exactly the input pixels, and the base64 text equals Python's
base64 of the PNG bytes. Rendered frames and int16 planes go through the
same path as the server uses them; the reference's fpng output of the same
pixels decodes to the same image."""
import base64
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_RENDER_VR, MkError
from tests.png_util import read_png

pytestmark = pytest.mark.gpu

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libmkref.so"


def _encode(img: np.ndarray):
    import torch

    from monkeygl_amd.encode import Encoder

    h, w, ch = img.shape
    enc = Encoder(w, h, ch)
    dev = torch.tensor(img, device="cuda")
    enc.encode_async(dev, w, h, ch, stream=torch.cuda.current_stream().cuda_stream)
    return enc.png_bytes(), enc.base64_text()


def _check(img):
    png, b64 = _encode(img)
    got, info = read_png(png)
    assert np.array_equal(got, img), (img.shape, np.argwhere(got != img)[:4])
    assert info["chunks"] == [b"IHDR", b"IDAT", b"IEND"]
    assert b64 == base64.b64encode(png).decode("ascii")
    return png, info


def _vr_like(h, w, seed=0):
    rng = np.random.default_rng(seed)
    img = np.zeros((h, w, 3), np.uint8)
    yy, xx = np.mgrid[0:h, 0:w]
    r = np.hypot(yy - h / 2, xx - w / 2)
    body = r < min(h, w) * 0.35
    img[body, 0] = (200 - r[body] / 2).astype(np.uint8)
    img[body, 1] = (xx[body] % 256).astype(np.uint8)
    img[body, 2] = rng.integers(0, 256, body.sum(), dtype=np.uint8)
    return img


@pytest.mark.parametrize("shape", [(1, 1, 3), (1, 1, 4), (2, 3, 3), (5, 7, 4), (33, 17, 3), (17, 33, 4),
                                   (64, 64, 3), (1, 4096, 3), (300, 1, 4), (8, 6144, 4), (3, 8192, 3)])
def test_random_and_edge_shapes(shape):
    rng = np.random.default_rng(sum(shape))
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    _check(img)


@pytest.mark.parametrize("value", [0, 143, 144, 255])
def test_constant_images_use_long_runs(value):
    # runs of every length class: a 4096-byte row is one literal and copies
    # of 258 plus a remainder; a handful of symbols makes a very short code
    img = np.full((6, 1365, 3), value, np.uint8)
    png, _ = _check(img)
    assert len(png) < img.size // 20


def test_run_lengths_across_windows():
    # runs ending at every offset within and across 64-byte windows
    rows = []
    for L in range(1, 300):
        row = np.zeros(1200, np.uint8)
        k = 0
        v = 0
        while k < 1200:
            row[k:k + L] = v
            v = (v + 37) % 256
            k += L
        rows.append(row)
    img = np.stack(rows).reshape(len(rows), 400, 3)
    _check(img)


def test_vr_like_frame_compresses_and_round_trips():
    img = _vr_like(512, 512)
    png, info = _check(img)
    assert len(png) < img.size   # the black background collapses
    assert set(info["filters"]) <= {0, 1, 2, 3, 4}


def test_rendered_frame_and_plane(gpu_ctx):
    import torch

    from monkeygl_amd.encode import Encoder, encode_plane_png_base64, encode_vr_png_base64

    vol = W.make_volume(W.SMALL)
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, P.transfer_function(W.tf_points(0)))
    info = P.RenderInfo()
    info.set_vr_size(256, 256)
    info.rotate(30.0, 10.0)
    p = P.render_params(W.SMALL.dims, W.SMALL.spacing, info, {0: P.LabelState(1.0, 400.0, 40.0)}, MK_RENDER_VR)
    dev = torch.empty(256 * 256 * 3, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    gpu_ctx.render_async(p, dev, stream=s)
    enc = Encoder(256, 256, 3)
    text = encode_vr_png_base64(enc, dev, 256, 256, stream=s)
    frame = dev.cpu().numpy().reshape(256, 256, 3)
    got, _ = read_png(base64.b64decode(text))
    assert np.array_equal(got, frame)
    # int16 plane: RGBA of width W/2 holds the plane's bytes
    q = P.plane_params(W.SMALL.dims, W.SMALL.spacing, P.PLANE_CORONAL)
    plane = torch.empty(q.width * q.height, dtype=torch.int16, device="cuda")
    gpu_ctx.render_plane_async(q, plane, stream=s)
    enc4 = Encoder(q.width // 2, q.height, 4)
    text = encode_plane_png_base64(enc4, plane, q.width, q.height, stream=s)
    got, info4 = read_png(base64.b64decode(text))
    assert info4["width"] == q.width // 2
    assert np.array_equal(got.reshape(-1).view(np.int16), plane.cpu().numpy())


@pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built")
def test_same_pixels_as_fpng():
    lib = C.CDLL(str(REF))
    lib.ref_fpng_encode.restype = C.c_ulonglong
    lib.ref_fpng_encode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_ulonglong]
    img = np.ascontiguousarray(_vr_like(300, 200, seed=4))
    out = np.zeros(img.size * 2 + 1024, np.uint8)
    n = lib.ref_fpng_encode(img.ctypes.data, 200, 300, 3, out.ctypes.data, out.size)
    ref, _ = read_png(out[:n].tobytes())
    png, _ = _encode(img)
    got, _ = read_png(png)
    assert np.array_equal(ref, got)


def test_argument_errors():
    import torch

    from monkeygl_amd import _capi

    lib = _capi.load()
    assert lib.mk_png_max_bytes(0, 4, 3) == 0
    assert lib.mk_png_max_bytes(4, 4, 2) == 0
    assert lib.mk_png_max_bytes(8193, 1, 4) == 0
    buf = torch.empty(1024, dtype=torch.uint8, device="cuda")
    st = lib.mk_png_encode_async(buf.data_ptr(), 16, 16, 3, buf.data_ptr(), 10, buf.data_ptr(), 10,
                                 buf.data_ptr(), 0)
    assert st != 0 and b"capacity" in lib.mk_last_error(None)


def _huffman_lengths(counts):
    """Textbook Huffman code lengths (unlimited), by repeated merging."""
    import heapq
    import itertools

    tick = itertools.count()
    heap = [(c, next(tick), [s]) for s, c in enumerate(counts) if c > 0]
    depth = {s: 0 for _, _, ss in heap for s in ss}
    heapq.heapify(heap)
    while len(heap) > 1:
        c1, _, s1 = heapq.heappop(heap)
        c2, _, s2 = heapq.heappop(heap)
        for s in s1 + s2:
            depth[s] += 1
        heapq.heappush(heap, (c1 + c2, next(tick), s1 + s2))
    return depth


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_huffman_code_is_complete_limited_and_optimal(seed):
    from tests.png_util import deflate_block_stats

    img = _vr_like(48, 64, seed=seed)
    png, _ = _check(img)
    ll, dl, counts = deflate_block_stats(png)
    assert max(ll) <= 12 and dl == [1]
    assert sum(2.0 ** -L for L in ll if L) == 1.0          # complete literal/length code
    assert all((c > 0) == (L > 0) for c, L in zip(counts, ll) if c > 0)
    cost = sum(c * L for c, L in zip(counts, ll))
    opt = _huffman_lengths(counts)
    if max(opt.values()) <= 12:                             # then the limit does not bind: Huffman's cost is the optimum
        assert cost == sum(counts[s] * d for s, d in opt.items())
    else:
        assert cost >= sum(counts[s] * d for s, d in opt.items())


def test_skewed_counts_hit_the_length_limit():
    from tests.png_util import deflate_block_stats

    # one byte value dominates, a few are rare: the unlimited Huffman code is deeper than 12
    img = np.zeros((64, 64, 4), np.uint8)
    rng = np.random.default_rng(9)
    pts = rng.integers(0, 64, (400, 2))
    img[pts[:, 0], pts[:, 1]] = rng.integers(1, 256, (400, 4), dtype=np.uint8)
    png, _ = _check(img)
    ll, _, counts = deflate_block_stats(png)
    assert max(ll) <= 12 and sum(2.0 ** -L for L in ll if L) == 1.0
