"""The pyMonkeyGL drop-in on the GPU, driven the way the reference's examples
drive pyMonkeyGL (examples/examples.py): its images against the one-request
entry points (bit-identical) and the FP64 oracle (1/255 on >= 99.9% of
pixels, trilinear MPR/CPR within 1 HU)."""
import base64

import numpy as np
import pytest

from monkeygl_amd import pyMonkeyGL as mk
from monkeygl_amd import workloads as W
from oracle import oracle as O
from tests.png_util import read_png

pytestmark = pytest.mark.gpu

SPEC = W.VolumeSpec("pmk", (96, 80, 72), (0.8, 0.8, 1.0), 6, (3, 12), (300, 1800), n_vessels=3, seed=41)


@pytest.fixture(scope="module")
def volume():
    return W.make_volume(SPEC)


def _hm(vol):
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(vol.swapaxes(2, 0))
    hm.SetSpacing(*SPEC.spacing)
    hm.SetTransferFunc({5: mk.RGBA(0.8, 0, 0, 0), 90: mk.RGBA(0.8, 0.8, 0.8, 0.8)})
    hm.SetVRWWWL(500, 100)
    return hm


def _vr_frame(hm):
    text = hm.GetVRData_pngString()
    img, info = read_png(base64.b64decode(text))
    assert info["channels"] == 3
    return img


def _luts(hm):
    from monkeygl_amd import params as P
    luts = {l: P.transfer_function(o.idx2rgba, o.idx2alpha) for l, o in hm._objects.items()}
    return {l: v for l, v in luts.items() if v is not None}


def _frac(a, b):
    d = np.abs(a.astype(int) - b.astype(int)).max(-1)
    return float((d <= 1).mean())


def test_vr_png_string_matches_plain_and_oracle(volume):
    hm = _hm(volume)
    hm.SetVRSize(200, 160)
    hm.Rotate(30, 20)
    img = _vr_frame(hm)
    p = hm._render_params(hm._ri.width, hm._ri.height)
    assert np.array_equal(img, hm._ctx.render(p))
    ref = O.render(p, volume, _luts(hm))
    assert _frac(img, ref) >= 0.999
    arr = hm.GetVRArray(200, 160)
    assert arr.shape == (3, 200, 160) and np.array_equal(arr.reshape(-1), img.reshape(-1))
    png = bytes(hm.GetVRData_png(200, 160))
    assert np.array_equal(read_png(png)[0], img)


def test_masks_labels_and_render_types(volume):
    hm = _hm(volume)
    hm.SetVRSize(128, 128)
    m = W.ellipsoid_mask(SPEC.dims, 2, seed=3)
    l1 = hm.AddNewObjectMaskArray((m == 1).astype(np.uint8).swapaxes(2, 0))
    l2 = hm.AddNewObjectMaskArray((m == 2).astype(np.uint8).swapaxes(2, 0))
    assert (l1, l2) == (1, 2)
    hm.SetTransferFunc({0: mk.RGBA(0.2, 0.9, 0.2, 0.1), 99: mk.RGBA(1, 1, 0, 0.9)}, l1)
    hm.SetVRWWWL(800, 300, l2)
    hm.SetObjectAlpha(0.4, 0)
    hm.Anterior()
    # the merged mask (the later object wins where they overlap), read back from the device
    merged = np.where(m == 2, 2, np.where(m == 1, 1, 0)).astype(np.uint8)
    assert np.array_equal(hm.GetMaskArray().reshape(-1), merged.reshape(-1))
    img = _vr_frame(hm)
    p = hm._render_params(128, 128)
    ref = O.render(p, volume, _luts(hm), mask_zyx=merged)
    assert _frac(img, ref) >= 0.999
    hm.SetRenderType(mk.RenderTypeMIP)
    mip = _vr_frame(hm)
    ref = O.render(hm._render_params(128, 128), volume, _luts(hm), mask_zyx=merged)
    assert _frac(mip, ref) >= 0.999


def test_planes_cpr_and_origin_slices(volume):
    hm = _hm(volume)
    hm.UpdateThickness(3.0)
    hm.SetMPRType(mk.MPRTypeMIP)
    hm.Browse(4.0, mk.PlaneAxial)
    hm.RotateCrossHair(25.0, mk.PlaneAxial)
    for plane in (mk.PlaneAxial, mk.PlaneSagittal, mk.PlaneCoronal):
        text = hm.GetPlaneData_pngString(plane)
        rgba, info = read_png(base64.b64decode(text))
        img, w, h = hm.GetPlaneData(plane)
        assert info["width"] * 2 == w and info["height"] == h
        assert np.array_equal(rgba.reshape(-1).view(np.int16).reshape(h, w), img)
        q = hm._mpr.plane_params(int(plane))
        ref = O.render_plane(q, volume)
        assert (np.abs(img.astype(int) - ref.astype(int)) <= 1).mean() >= 0.999, plane
    line = np.stack([np.linspace(20, 70, 60), np.full(60, 40.0), np.linspace(10, 60, 60)], 1).astype(np.float32)
    assert hm.SetCPRLineVoxelArray(line)
    hm.RotateCPR(30.0, mk.PlaneStretchedCPR)
    from monkeygl_amd import cpr as G
    for plane, which in ((mk.PlaneStretchedCPR, G.PLANE_STRETCHED_CPR),
                         (mk.PlaneStraightenedCPR, G.PLANE_STRAIGHTENED_CPR)):
        rgba, _ = read_png(base64.b64decode(hm.GetPlaneData_pngString(plane)))
        pts, dirs, w, h = hm._cprinfo.get_cpr_info(which)
        got = rgba.reshape(-1).view(np.int16).reshape(h, w)
        ref = O.render_cpr(w, pts, dirs, False, volume, 0)
        assert (np.abs(got.astype(int) - ref.astype(int)) <= 1).mean() >= 0.999, plane
    # an original slice, -2048 outside the mask once there is one
    rgba, _ = read_png(base64.b64decode(hm.GetOriginData_pngString(30)))
    assert np.array_equal(rgba.reshape(-1).view(np.int16), volume[30].reshape(-1))
    m = np.zeros(volume.shape, np.uint8)
    m[:, 10:40, 10:40] = 1
    hm.AddNewObjectMaskArray(m.swapaxes(2, 0))
    rgba, _ = read_png(base64.b64decode(hm.GetOriginData_pngString(30)))
    exp = np.where(m[30] > 0, volume[30], -2048).astype(np.int16)
    assert np.array_equal(rgba.reshape(-1).view(np.int16), exp.reshape(-1))


def test_two_instances_are_independent(volume):
    a = _hm(volume)
    b = _hm(np.ascontiguousarray(volume[::-1]))
    for hm in (a, b):
        hm.SetVRSize(96, 96)
    b.SetVRWWWL(300, 0)
    fa1 = _vr_frame(a)
    fb = _vr_frame(b)
    fa2 = _vr_frame(a)
    assert np.array_equal(fa1, fa2) and not np.array_equal(fa1, fb)


def test_annotations_drawn_in_vr_frames(volume):
    """AddAnnotation's texts on the VR frame (core/AnnotationLayer.cpp:41-61,
    core/VRProvider.cpp:30-36): the frame equals the plain render with the
    oracle's Textout2Image of each annotation applied in order; removed
    annotations and a disabled layer leave the plain render."""
    hm = _hm(volume)
    hm.SetVRSize(160, 120)
    hm.Rotate(15, 10)
    rng = np.random.default_rng(4)
    glyphs = {}
    for font, h in ((0, 22), (2, 44)):
        for ch in "HiOK!0123456789 ":
            w = int(rng.integers(3, 14))
            m = rng.integers(0, 256, size=(h, w)).astype(np.uint8)
            m[rng.random((h, w)) < 0.3] = 0
            glyphs[(font, ord(ch))] = m
            hm.SetAnnotationGlyph(font, ch, m)
    annos = [("Hi 42", 4, 3, mk.FontSizeSmall, mk.AnnotationFormatLeft, mk.RGBA(1.0, 1.0, 0.0)),
             ("OK!", 156, 70, mk.FontSizeBig, mk.AnnotationFormatRight, mk.RGBA(0.2, 0.9, 0.5)),
             ("W?9", 80, 100, mk.FontSizeSmall, mk.AnnotationFormatCenter, mk.RGBA(0.5, 0.5, 1.0))]
    for a in annos:
        assert hm.AddAnnotation(mk.PlaneVR, *a)
    img = _vr_frame(hm)
    p = hm._render_params(hm._ri.width, hm._ri.height)
    plain = hm._ctx.render(p)
    exp = plain.copy()
    table = O.glyph_table(glyphs)
    for txt, x, y, fs, fmt, clr in annos:
        O.textout(exp, txt.encode(), x, y, int(fs), int(fmt), clr.tuple()[:3], table)
    assert not np.array_equal(exp, plain)
    assert np.array_equal(img, exp)
    hm.EnableLayer(mk.LayerTypeAnnotation, False)
    assert np.array_equal(_vr_frame(hm), plain)
    hm.EnableLayer(mk.LayerTypeAnnotation, True)
    hm.RemoveAllAnnotations()
    assert np.array_equal(_vr_frame(hm), plain)
