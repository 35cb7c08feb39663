"""Annotation text on the GPU (mk_set_glyph + mk_draw_text_async) against the
oracle's restatement of AnnotationUtils::Textout2Image, bit for bit.

Glyphs are synthetic (seeded masks with zero and nonzero bytes, three font
sizes of different heights, a subset of the bytes), so bytes without a glyph
occur. Annotations of every format and font, placed partly or wholly outside
the image, with colours in [0, 1] and beyond, are drawn one after another
into one image, as the annotation layer draws its list; a text longer than
one launch's glyphs is included. Glyphs replaced or removed between draws
take effect at the next draw.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _glyphs(rng):
    g = {}
    for font, h in ((0, 22), (1, 33), (2, 44)):
        for code in rng.choice(np.arange(32, 127), size=60, replace=False):
            w = int(rng.integers(1, 20))
            m = rng.integers(0, 256, size=(h, w)).astype(np.uint8)
            m[rng.random((h, w)) < 0.4] = 0
            g[(font, int(code))] = m
    return g


def _annotations(rng, n, width, height):
    out = []
    for _ in range(n):
        k = int(rng.integers(1, 24))
        text = bytes(rng.integers(32, 127, size=k).astype(np.uint8))
        x = int(rng.integers(-60, width + 20))
        y = int(rng.integers(-30, height + 10))
        font = int(rng.choice([0, 1, 2, 2, -1]))
        fmt = int(rng.choice([0, 1, 2, -1]))
        color = rng.uniform(0.0, 1.0, 3) if rng.random() < 0.8 else rng.uniform(-1.0, 3.0, 3)
        out.append((text, x, y, font, fmt, color.astype(np.float32)))
    return out


def _draw_both(ctx, glyphs, annos, img0, stream):
    table = O.glyph_table(glyphs)
    h, w = img0.shape[:2]
    dev = torch.tensor(img0.reshape(-1), device="cuda")
    ref = img0.copy()
    for text, x, y, font, fmt, color in annos:
        ctx.draw_text_async(text, x, y, font, fmt, color, dev, w, h, stream)
        O.textout(ref, text, x, y, font, fmt, color, table)
    torch.cuda.synchronize()
    return dev.cpu().numpy().reshape(h, w, 3), ref


@pytest.mark.parametrize("size", [(320, 240), (97, 61), (10, 10)])
def test_text_matches_textout2image(gpu_ctx, size):
    rng = np.random.default_rng(size[0])
    glyphs = _glyphs(rng)
    for (font, code), m in glyphs.items():
        gpu_ctx.set_glyph(font, code, m)
    w, h = size
    img0 = rng.integers(0, 256, size=(h, w, 3)).astype(np.uint8)
    annos = _annotations(rng, 40, w, h)
    # one text longer than a launch's glyphs, through the middle
    long_text = bytes(rng.integers(32, 127, size=300).astype(np.uint8))
    annos.append((long_text, -500, h // 3, 0, 0, np.array([0.9, 0.4, 0.7], np.float32)))
    stream = torch.cuda.current_stream().cuda_stream
    got, ref = _draw_both(gpu_ctx, glyphs, annos, img0, stream)
    changed = int((ref != img0).any(-1).sum())
    print(f"{w}x{h}: {len(annos)} annotations, {changed} pixels changed")
    assert np.array_equal(got, ref)
    if min(w, h) >= 10 and w > 10:
        assert changed > 0
    # replace one glyph, remove another: the next draw uses the new set
    (f0, c0), (f1, c1) = list(glyphs)[:2]
    glyphs[(f0, c0)] = np.full((7, 3), 99, np.uint8)
    del glyphs[(f1, c1)]
    gpu_ctx.set_glyph(f0, c0, glyphs[(f0, c0)])
    gpu_ctx.set_glyph(f1, c1, None)
    t = bytes([c0, c1, c0])
    got2, ref2 = _draw_both(gpu_ctx, glyphs, [(t, 3, 2, f0, 0, np.ones(3, np.float32)),
                                              (t, 3, 30, f1, 2, np.ones(3, np.float32))], img0, stream)
    assert np.array_equal(got2, ref2)
    for (font, code) in glyphs:
        gpu_ctx.set_glyph(font, code, None)


def test_text_arguments_are_checked(gpu_ctx):
    from monkeygl_amd._capi import MkError

    dev = torch.zeros(64 * 64 * 3, dtype=torch.uint8, device="cuda")
    with pytest.raises(MkError):
        gpu_ctx.draw_text_async(b"a", 0, 0, 3, 0, (1, 1, 1), dev, 64, 64)      # no such font size
    with pytest.raises(MkError):
        gpu_ctx.draw_text_async(b"a", 0, 0, 0, 5, (1, 1, 1), dev, 64, 64)      # no such format
    with pytest.raises(MkError):
        gpu_ctx.set_glyph(0, 65, np.zeros((2000, 3), np.uint8))                 # taller than 1024
    with pytest.raises(MkError):
        gpu_ctx.set_glyph(4, 65, np.zeros((3, 3), np.uint8))
