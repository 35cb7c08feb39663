"""Pins of the oracle's Textout2Image restatement (no GPU).

AnnotationUtils::Textout2Image / GetSize (core/AnnotationUtils.cpp:419-475),
restated in oracle/mk_oracle.cpp over a glyph table, against closed forms:
placement of left, right and centred text, the default mask of a byte without
a glyph (width -1: the cursor moves one pixel left), overlap and mask-zero
pixels, clipping at the image edges, the colour conversion, and the early
returns.
"""
import numpy as np

from oracle import oracle as O

LEFT, RIGHT, CENTER, NONE = 0, 1, 2, -1


def _img(w=64, h=48, fill=0):
    return np.full((h, w, 3), fill, np.uint8)


def _table(glyphs):
    return O.glyph_table(glyphs)


def test_full_glyph_fills_its_box_with_the_colour():
    t = _table({(0, ord("a")): np.full((6, 4), 255, np.uint8)})
    img = _img()
    assert O.textout(img, b"a", 20, 15, 0, LEFT, (1.0, 0.5, 0.25), t)
    box = img[15:21, 20:24]
    assert np.all(box == np.array([255, 127, 63], np.uint8))    # 255*0.5 = 127.5, 255*0.25 = 63.75
    rest = img.copy()
    rest[15:21, 20:24] = 0
    assert not rest.any()


def test_right_and_centre_placement():
    g = {(1, ord("a")): np.full((3, 5), 9, np.uint8), (1, ord("b")): np.full((3, 8), 9, np.uint8)}
    t = _table(g)
    for fmt, x0 in ((LEFT, 30), (RIGHT, 30 - 13), (CENTER, 30 - 6)):   # width 13, 13 / 2 = 6
        img = _img()
        O.textout(img, b"ab", 30, 10, 1, fmt, (1, 1, 1), t)
        cols = np.flatnonzero(img[10, :, 0])
        assert cols[0] == x0 and cols[-1] == x0 + 12, (fmt, cols)


def test_byte_without_glyph_moves_the_cursor_left_and_the_later_glyph_wins():
    a = np.full((4, 5), 100, np.uint8)
    b = np.zeros((4, 3), np.uint8)
    b[:, 0] = 200   # b's first column overlaps a's last one
    t = _table({(0, ord("a")): a, (0, ord("b")): b})
    img = _img()
    O.textout(img, b"a?b", 10, 5, 0, LEFT, (1, 1, 1), t)
    assert np.all(img[5:9, 10:14, 0] == 100)
    assert np.all(img[5:9, 14, 0] == 200)          # '?' moved the cursor from 15 back to 14
    assert not img[5:9, 15:17].any()               # b's zero columns leave the image alone


def test_centre_divides_a_negative_width_toward_zero():
    a = np.full((2, 2), 50, np.uint8)
    t = _table({(0, ord("a")): a})
    img = _img()
    # width -1 -1 -1 + 2 = -1; -1 / 2 == 0 in C++ (floor would give -1)
    O.textout(img, b"???a", 20, 5, 0, CENTER, (1, 1, 1), t)
    cols = np.flatnonzero(img[5, :, 0])
    assert list(cols) == [17, 18]


def test_mask_zero_pixels_keep_the_image():
    m = np.array([[0, 7], [7, 0]], np.uint8)
    t = _table({(2, ord("x")): m})
    img = _img(fill=33)
    O.textout(img, b"x", 3, 4, 2, LEFT, (1, 1, 1), t)
    assert img[4, 3, 0] == 33 and img[4, 4, 0] == 7 and img[5, 3, 0] == 7 and img[5, 4, 0] == 33


def test_clipped_at_every_edge():
    m = np.arange(1, 1 + 10 * 12, dtype=np.uint8).reshape(10, 12)
    t = _table({(0, ord("q")): m})
    for x, y in ((-5, -4), (58, 42), (-11, 20), (30, -9)):
        img = _img()
        O.textout(img, b"q", x, y, 0, LEFT, (1, 1, 1), t)
        exp = _img()
        for gy in range(10):
            for gx in range(12):
                xx, yy = x + gx, y + gy
                if 0 <= xx < 64 and 0 <= yy < 48:
                    exp[yy, xx] = m[gy, gx]
        assert np.array_equal(img, exp), (x, y)


def test_colour_products_are_truncated_and_kept_to_8_bits():
    t = _table({(0, ord("c")): np.full((1, 1), 200, np.uint8)})
    img = _img()
    O.textout(img, b"c", 0, 0, 0, LEFT, (2.0, 0.999, 0.0), t)
    assert list(img[0, 0]) == [400 % 256, 199, 0]   # 400 -> 144; 199.8 -> 199


def test_early_returns_draw_nothing():
    t = _table({(0, ord("a")): np.full((3, 3), 255, np.uint8)})
    for w, h in ((9, 48), (64, 9)):
        img = _img(w, h)
        assert not O.textout(img, b"a", 0, 0, 0, LEFT, (1, 1, 1), t) and not img.any()
    img = _img()
    assert not O.textout(img, b"a", 0, 0, 0, NONE, (1, 1, 1), t) and not img.any()
    assert O.textout(img, b"a", 0, 0, -1, LEFT, (1, 1, 1), t) and not img.any()   # no font: default masks
