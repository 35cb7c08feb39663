"""The orientation figure on the GPU (mk_merge_orientation_box_async): the
whole image after the merge equals the reference's own rasteriser
(Methods::FillHoleInImage_Ch1 under MergeOrientationBox, built into
oracle/_ref) on the same input, bit for bit, for several rotations, image
sizes and figures; the pyMonkeyGL drop-in draws the figure when
./resources/human.stl exists, as the reference's MarchingCube does."""
import base64
import ctypes as C
import struct
from pathlib import Path

import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd.workloads import figure, stl_bytes, uv_ellipsoid

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libmkref.so"
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built")]


def _ref_box(stl, rot, img):
    L = C.CDLL(str(REF))
    L.ref_stl_normalize.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.ref_orientation_box.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    n = struct.unpack_from("<I", stl, 80)[0]
    body = np.frombuffer(stl, np.uint8, 50 * n, 84).copy()
    fac = np.zeros(n * 9, np.float32)
    L.ref_stl_normalize(body.ctypes.data, n, fac.ctypes.data)
    out = img.copy()
    r = np.ascontiguousarray(rot, np.float32)
    L.ref_orientation_box(fac.ctypes.data, n, r.ctypes.data, out.ctypes.data, img.shape[1], img.shape[0])
    return out


def _rot(ax, ay, zoom=1.0):
    info = P.RenderInfo()
    info.rotate(ax, ay)
    if zoom != 1.0:
        info.zoom(zoom)
    return np.asarray(info.rot, np.float32)


@pytest.mark.parametrize("seed", [0, 3])
@pytest.mark.parametrize("angles", [(0, 0), (30, 20), (135, -40), (250, 75), (90, 0), (0, 90)])
def test_matches_the_reference_rasteriser(gpu_ctx, seed, angles):
    import torch

    stl = stl_bytes(figure(seed))
    assert gpu_ctx.set_orientation_mesh(stl) == struct.unpack_from("<I", stl, 80)[0]
    ws = torch.empty(gpu_ctx.orientation_workspace_bytes(), dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(seed)
    s = torch.cuda.current_stream().cuda_stream
    for w, h in ((100, 90), (513, 300), (65, 65), (64, 200), (200, 64)):
        img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        dev = torch.tensor(img, device="cuda")
        gpu_ctx.merge_orientation_box_async(_rot(*angles), dev, w, h, ws, stream=s)
        got = dev.cpu().numpy()
        exp = _ref_box(stl, _rot(*angles), img)
        assert np.array_equal(got, exp), (w, h, int((got != exp).any(-1).sum()))
    gpu_ctx.set_orientation_mesh(None)


def test_fine_figure_and_removal(gpu_ctx):
    import torch

    # a finely tessellated figure (tens of thousands of small facets, like the reference's)
    parts = [uv_ellipsoid((0, 0, 100), (25, 15, 60), 140, 90), uv_ellipsoid((0, 0, 175), (12, 12, 15), 90, 60)]
    stl = stl_bytes(np.concatenate(parts))
    gpu_ctx.set_orientation_mesh(stl)
    ws = torch.empty(gpu_ctx.orientation_workspace_bytes(), dtype=torch.uint8, device="cuda")
    img = np.zeros((256, 256, 3), np.uint8)
    dev = torch.tensor(img, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    gpu_ctx.merge_orientation_box_async(_rot(40, 10), dev, 256, 256, ws, stream=s)
    got = dev.cpu().numpy()
    assert (got.sum(-1) > 0).sum() > 100
    assert np.array_equal(got, _ref_box(stl, _rot(40, 10), img))
    gpu_ctx.set_orientation_mesh(None)
    dev.zero_()
    gpu_ctx.merge_orientation_box_async(_rot(40, 10), dev, 256, 256, ws, stream=s)
    assert int(dev.sum().item()) == 0


def test_drop_in_draws_the_figure(tmp_path, monkeypatch):
    from monkeygl_amd import pyMonkeyGL as mk
    from monkeygl_amd import workloads as W
    from tests.png_util import read_png

    stl = stl_bytes(figure(1))
    (tmp_path / "resources").mkdir()
    (tmp_path / "resources" / "human.stl").write_bytes(stl)
    monkeypatch.chdir(tmp_path)
    spec = W.VolumeSpec("fig", (48, 40, 36), (1.0, 1.0, 1.0), 3, (2, 6), (300, 1500), n_vessels=1, seed=4)
    vol = W.make_volume(spec)
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(vol.swapaxes(2, 0))
    hm.SetVRWWWL(500, 100)
    hm.SetVRSize(160, 120)
    hm.Rotate(30, 20)
    img, _ = read_png(base64.b64decode(hm.GetVRData_pngString()))
    plain = hm._ctx.render(hm._render_params(160, 120))
    exp = _ref_box(stl, np.asarray(hm._ri.rot, np.float32), plain)
    assert not np.array_equal(exp, plain) and np.array_equal(img, exp)
    arr = hm.GetVRArray(160, 120)
    assert np.array_equal(arr.reshape(-1), exp.reshape(-1))


def _ref_polyline(xy, rgb, img):
    L = C.CDLL(str(REF))
    L.ref_polyline_24bit.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    out = img.copy()
    a = np.ascontiguousarray(xy, np.float32)
    c = np.ascontiguousarray(rgb, np.float32)
    L.ref_polyline_24bit(a.ctypes.data, a.shape[0], c.ctypes.data, out.ctypes.data, img.shape[1], img.shape[0])
    return out


def test_polyline_matches_the_reference_line_drawer(gpu_ctx):
    import torch

    rng = np.random.default_rng(11)
    h, w = 150, 210
    img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    xy = np.concatenate([rng.uniform(-40, 250, (30, 2)),
                         [[10.5, 10.5], [10.5, 10.5], [199.99, 149.2], [-5.0, 160.0], [0.0, 0.0]]]).astype(np.float32)
    for color in ((1.0, 1.0, 1.0), (0.9, 0.25, 0.0)):
        dev = torch.tensor(img, device="cuda")
        gpu_ctx.draw_polyline_async(torch.tensor(xy, device="cuda"), len(xy), color, dev, w, h,
                                    stream=torch.cuda.current_stream().cuda_stream)
        assert np.array_equal(dev.cpu().numpy(), _ref_polyline(xy, color, img)), color


def test_drop_in_draws_the_cpr_line():
    from monkeygl_amd import pyMonkeyGL as mk
    from monkeygl_amd import workloads as W
    from tests.png_util import read_png

    spec = W.VolumeSpec("cprline", (48, 40, 36), (0.8, 0.8, 1.2), 3, (2, 6), (300, 1500), n_vessels=1, seed=6)
    vol = W.make_volume(spec)
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(vol.swapaxes(2, 0))
    hm.SetSpacing(*spec.spacing)
    hm.SetVRWWWL(500, 100)
    hm.SetVRSize(180, 140)
    hm.Rotate(25, 15)
    line = np.stack([np.linspace(5, 40, 12), 20 + 8 * np.sin(np.linspace(0, 3, 12)), np.linspace(4, 30, 12)], 1)
    assert hm.SetCPRLineVoxelArray(line.astype(np.float32))
    hm.SetCPRLineColor(mk.RGBA(1.0, 0.2, 0.1))
    img, _ = read_png(base64.b64decode(hm.GetVRData_pngString()))
    plain = hm._ctx.render(hm._render_params(180, 140))
    xy = hm._voxel_to_vr(hm._cprinfo.line, 180, 140)
    exp = _ref_polyline(xy, (1.0, 0.2, 0.1), plain)
    assert not np.array_equal(exp, plain) and np.array_equal(img, exp)
    hm.EnableLayer(mk.LayerTypeCPRLine, False)
    img, _ = read_png(base64.b64decode(hm.GetVRData_pngString()))
    assert np.array_equal(img, plain)
