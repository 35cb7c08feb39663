"""Host behaviour of the pyMonkeyGL drop-in (no GPU): names and enum values of
the reference module, the raw-memory reading of arrays, object/label
bookkeeping, the alpha/window table rule, TF files and the render arguments."""
import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import pyMonkeyGL as mk
from monkeygl_amd._capi import MK_MASK_ADD, MK_MASK_UPDATE


def test_enum_values_and_exports():
    # core/Defines.h:151-262
    assert mk.PlaneAxial == 0 and mk.PlaneCoronalOblique == 5 and mk.PlaneVR == 6
    assert mk.PlaneStretchedCPR == 7 and mk.PlaneStraightenedCPR == 8 and mk.PlaneNotDefined == -1
    assert mk.MPRTypeAverage == 0 and mk.MPRTypeMinIP == 2
    assert mk.RenderTypeVR == 0 and mk.RenderTypeSurface == 2
    assert mk.LayerTypeCPRLine == 2 and mk.LogLevelError == 2 and mk.OrientationFoot == 5
    assert mk.FontSizeBig == 2 and mk.AnnotationFormatCenter == 2
    for name in ("SetVolumeArray", "AddNewObjectMaskArray", "UpdateMaskArray", "SetSpacing", "SetDirection",
                 "SetOrigin", "SetTransferFunc", "LoadTransferFunction", "SaveTransferFunction",
                 "SetColorBackground", "Reset", "Anterior", "Posterior", "Left", "Right", "Head", "Foot",
                 "SetVRWWWL", "SetObjectAlpha", "SetRenderType", "Rotate", "Pan", "Zoom", "UpdateThickness",
                 "SetThickness", "GetThickness", "SetMPRType", "Browse", "PanCrossHair", "RotateCrossHair",
                 "SetCPRLinePatientArray", "SetCPRLineVoxelArray", "RotateCPR", "SetCPRLineColor", "SetVRSize",
                 "ShowPlaneInVR", "SetPlaneIndex", "Transfer2Base64Array", "AddAnnotation",
                 "RemovePlaneAnnotations", "RemoveAllAnnotations", "EnableLayer", "GetZoomRatio",
                 "GetPlaneCurrentIndex", "GetPlaneTotalNumber", "GetVolumeArray", "GetVRArray",
                 "GetVRData_pngString", "GetVRData_png", "SaveVR2Png", "GetPlaneData_pngString",
                 "GetOriginData_pngString", "GetCrossHairPoint", "GetPlaneIndex", "GetPlaneNumber",
                 "SetLogLevel", "LoadVolumeFile", "AddObjectMaskFile"):
        assert callable(getattr(mk.HelloMonkey, name)), name


def test_value_types():
    c = mk.RGBA(0.1, 0.2, 0.3)
    assert c.tuple() == (0.1, 0.2, 0.3, 1.0)
    assert mk.RGBA().tuple() == (0.0, 0.0, 0.0, 0.0)
    d = mk.Direction3d(0, 3, 4)
    assert np.allclose(d.tuple(), (0, 0.6, 0.8))
    assert mk.Direction3d().tuple() == (1.0, 0.0, 0.0)
    p = mk.Point3d(1, 2, 3)
    p.SetZ(5)
    assert (p.x(), p.y(), p.z()) == (1, 2, 5)


def _vol(W=24, H=20, D=16, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(-1000, 2000, (D, H, W), dtype=np.int16)


def test_volume_read_as_raw_memory():
    zyx = _vol()
    hm = mk.HelloMonkey()
    assert hm.SetVolumeArray(zyx.swapaxes(2, 0))
    assert hm._dims == (24, 20, 16)
    assert np.array_equal(hm._vol, zyx.reshape(-1))
    # GetVolumeArray hands the raw buffer back shaped (W, H, D)
    assert np.array_equal(hm.GetVolumeArray().reshape(-1), zyx.reshape(-1))


class _Recorder:
    """Stands in for the device context: records the mask merges the
    module asks for (the merge itself runs on the GPU and is tested there)."""

    def __init__(self):
        self.merges = []

    def merge_mask(self, arr, label, mode):
        self.merges.append((arr.copy(), label, mode))


def _recorded(hm):
    rec = _Recorder()
    hm._device = lambda: rec
    return rec


def test_label_allocation_and_mask_merges():
    zyx = _vol()
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(zyx.swapaxes(2, 0))
    rec = _recorded(hm)
    m1 = np.zeros_like(zyx, dtype=np.uint8)
    m1[2:6, 3:9, 4:10] = 1
    m2 = np.zeros_like(m1)
    m2[5:9, 5:12, 8:14] = 7
    hm.SetVRWWWL(300, 20)
    assert not hm.UpdateMaskArray(m1.swapaxes(2, 0), 0)    # no mask yet
    assert hm.AddNewObjectMaskArray(m1.swapaxes(2, 0)) == 1
    # the new label copies the active object's settings
    assert hm._objects[1].ww == np.float32(300)
    assert hm.AddNewObjectMaskArray(m2.swapaxes(2, 0)) == 2
    # the (W, H, D) views reach the device as the x-fastest buffers, add mode
    assert [(l, md) for _, l, md in rec.merges] == [(1, MK_MASK_ADD), (2, MK_MASK_ADD)]
    assert np.array_equal(rec.merges[0][0], m1) and np.array_equal(rec.merges[1][0], m2)
    m3 = np.zeros_like(m1)
    m3[0:2, 0:2, 0:2] = 1
    assert hm.UpdateMaskArray(m3.swapaxes(2, 0), 1)
    assert rec.merges[-1][1:] == (1, MK_MASK_UPDATE) and np.array_equal(rec.merges[-1][0], m3)
    assert not hm.UpdateMaskArray(m3.swapaxes(2, 0), 9)   # no such object
    assert not hm.UpdateMaskArray(m3[:2].swapaxes(2, 0), 1)   # smaller than the volume
    assert hm.AddNewObjectMaskArray(m3[:2].swapaxes(2, 0)) == 0
    # a gap in the labels is filled first
    del hm._objects[1]
    assert hm.AddNewObjectMaskArray(m3.swapaxes(2, 0)) == 1
    # a new volume drops the mask and the objects
    hm.SetVolumeArray(zyx.swapaxes(2, 0))
    assert not hm._has_mask and list(hm._objects) == [0]


def test_label_capacity():
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(_vol(8, 8, 8).swapaxes(2, 0))
    _recorded(hm)
    one = np.ones((8, 8, 8), np.uint8)
    labels = [hm.AddNewObjectMaskArray(one) for _ in range(65)]
    assert labels[:64] == list(range(1, 65)) and labels[64] == 0


def test_alpha_window_table_rule():
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(_vol().swapaxes(2, 0))
    # no SetVRWWWL / SetObjectAlpha yet: the table keeps (0, 0, 0)
    assert hm._aw[0] == (0.0, 0.0, 0.0)
    assert hm.SetVRWWWL(-5, 40)            # ww <= 0 becomes 1
    assert hm._aw[0] == (1.0, 1.0, 40.0)
    assert hm.SetObjectAlpha(0.4)
    assert hm._aw[0] == (np.float32(0.4), 1.0, 40.0)
    assert not hm.SetVRWWWL(400, 40, 65)
    # SetTransferFunc needs an existing object
    assert not hm.SetTransferFunc({0: mk.RGBA(1, 0, 0, 1), 99: mk.RGBA(0, 1, 0, 1)}, 3)
    assert hm.SetTransferFunc({0: mk.RGBA(1, 0, 0, 1), 99: mk.RGBA(0, 1, 0, 1)}, {10: 0.2, 80: 0.9})
    assert hm._objects[0].idx2alpha == {10: 0.2, 80: 0.9}


def test_transfer_function_file_round_trip(tmp_path):
    # trfns/coro.txt of the reference, as data
    f = tmp_path / "coro.txt"
    f.write_text("ww: 500\nwl: 100\nidx2rgba: 2\n5 0.8 0 0 0 \n90 0.8 0.8 0.8 0.8 \nidx2alpha: 0\n")
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(_vol().swapaxes(2, 0))
    assert hm.LoadTransferFunction(str(f))
    info = hm._objects[0]
    assert (info.ww, info.wl) == (500.0, 100.0)
    assert info.idx2rgba == {5: (np.float32(0.8), 0.0, 0.0, 0.0), 90: (np.float32(0.8),) * 4}
    g = tmp_path / "out.txt"
    assert hm.SaveTransferFunction(str(g))
    assert g.read_text().split() == f.read_text().split()


def test_invert_z_and_render_arguments():
    zyx = _vol()
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(zyx.swapaxes(2, 0))
    hm.SetSpacing(0.7, 0.8, 1.2)
    hm.SetVRWWWL(400, 40)
    assert not hm._need_invert_z()   # default directions: z = (0, 0, -1)
    hm.SetDirection(mk.Direction3d(1, 0, 0), mk.Direction3d(0, 1, 0), mk.Direction3d(0, 0, 1))
    assert hm._need_invert_z()
    hm.SetVRSize(300, 200)
    hm.Rotate(30, 15)
    p = hm._render_params(hm._ri.width, hm._ri.height)
    info = P.RenderInfo()
    info.set_vr_size(300, 200)
    info.rotate(30, 15)
    q = P.render_params((24, 20, 16), (0.7, 0.8, 1.2), info, {0: P.LabelState(1.0, 400.0, 40.0)},
                        invert_z=True, background=(0, 0, 0, 0))
    assert bytes(p) == bytes(q)


def test_mpr_and_cpr_state_through_the_api():
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(_vol(64, 64, 48).swapaxes(2, 0))
    hm.SetSpacing(1.0, 1.0, 1.5)
    n0 = hm.GetPlaneNumber(mk.PlaneAxial)
    assert n0 == int(48 * 1.5 / 1.0)
    hm.SetPlaneIndex(10, mk.PlaneAxial)
    assert hm.GetPlaneCurrentIndex(mk.PlaneAxial) in (9, 10)
    hm.Browse(2.0, mk.PlaneAxial)
    pt = hm.GetCrossHairPoint(mk.PlaneSagittal)
    assert isinstance(pt, mk.Point2d)
    line = np.stack([np.full(30, 32.0), np.full(30, 30.0), np.linspace(5, 40, 30)], 1).astype(np.float32)
    assert hm.SetCPRLineVoxelArray(line)
    assert hm.RotateCPR(15.0, mk.PlaneStretchedCPR)
    assert not hm.RotateCPR(15.0, mk.PlaneAxial)
    assert not hm.SetCPRLineVoxelArray(np.zeros((5, 2), np.float32))


def test_file_readers_are_reported_missing():
    hm = mk.HelloMonkey()
    with pytest.raises(NotImplementedError):
        hm.LoadVolumeFile("x.mhd")
