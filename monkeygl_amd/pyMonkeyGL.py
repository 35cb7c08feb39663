"""Drop-in for the reference's Python module ``pyMonkeyGL``.

The reference exposes ``HelloMonkey`` and its value types through pybind11
(pybind11_interface/PybindInterface.cpp:186-351 over core/HelloMonkey.h:36-139
and core/DataManager.cpp). This module keeps those names, overloads and
argument meanings, and renders through the HIP C ABI (include/monkeygl_hip.h):

    from monkeygl_amd import pyMonkeyGL as mk
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(np_zyx.swapaxes(2, 0))
    hm.SetSpacing(0.7, 0.7, 1.0)
    hm.SetTransferFunc({5: mk.RGBA(0.8, 0, 0, 0), 90: mk.RGBA(0.8, 0.8, 0.8, 0.8)})
    hm.SetVRWWWL(500, 100)
    text = hm.GetVRData_pngString()

State that the reference keeps in one process-wide DataManager
(core/DataManager.cpp:46-54) lives in each HelloMonkey here, with its own
device context: two instances render two volumes independently. Device data
(volume, mask, LUTs) is uploaded when a render first needs it.

Kept as the reference has them: the volume and mask arrays are read as their
raw memory, x fastest, with dims taken from shape[0..2] (callers pass
``np_zyx.swapaxes(2, 0)`` views, PybindInterface.cpp:65-78); the label
alpha/window table is refreshed only by SetVRWWWL / SetObjectAlpha
(DataManager.cpp:565-611), so labels never given a window keep (0, 0, 0);
GetVRArray returns the frame's bytes shaped (3, w, h); GetVRData_png returns
a list of byte values (pybind11's std::vector<uint8_t>).

Differences, by design: images come back as PNGs built on the device (same
pixels as fpng's, other bytes); the mask of a first AddNewObjectMaskArray is
zero outside the object (the reference leaves that memory uninitialised,
VolumeInfo.cpp:100-107); the VR overlays are drawn on the device in the
reference's layer order (the orientation figure from ./resources/human.stl,
then the annotation text, then the CPR line), and the text's glyph masks are
data (./resources/annotation_glyphs.npz or SetAnnotationGlyph), since the
reference compiles its own into its library; LoadVolumeFile and
AddObjectMaskFile need ITK readers, which are not part of this build;
SaveVR2Png writes the PNG (the reference writes width*height*3 bytes of it).
"""
from __future__ import annotations

import base64 as _b64
import copy
import enum
import math

import numpy as np

from . import cpr as _cpr
from . import mpr as _mpr
from . import params as _P
from ._capi import (MK_MASK_ADD, MK_MASK_UPDATE, MK_NUM_LABELS, MK_SAMPLE_LINEAR, MkError)

MAXOBJECTCOUNT = 64


# ---- enums (core/Defines.h:151-262), exported at module level --------------

class LogLevel(enum.IntEnum):
    LogLevelNotDefined = -1
    LogLevelInfo = 0
    LogLevelWarn = 1
    LogLevelError = 2


class PlaneType(enum.IntEnum):
    PlaneNotDefined = -1
    PlaneAxial = 0
    PlaneSagittal = 1
    PlaneCoronal = 2
    PlaneAxialOblique = 3
    PlaneSagittalOblique = 4
    PlaneCoronalOblique = 5
    PlaneVR = 6
    PlaneStretchedCPR = 7
    PlaneStraightenedCPR = 8


class Orientation(enum.IntEnum):
    OrientationNotDefined = -1
    OrientationAnterior = 0
    OrientationPosterior = 1
    OrientationLeft = 2
    OrientationRight = 3
    OrientationHead = 4
    OrientationFoot = 5


class MPRType(enum.IntEnum):
    MPRTypeNotDefined = -1
    MPRTypeAverage = 0
    MPRTypeMIP = 1
    MPRTypeMinIP = 2


class RenderType(enum.IntEnum):
    RenderTypeNotDefined = -1
    RenderTypeVR = 0
    RenderTypeMIP = 1
    RenderTypeSurface = 2


class LayerType(enum.IntEnum):
    LayerTypeNotDefined = -1
    LayerTypeImage = 0
    LayerTypeAnnotation = 1
    LayerTypeCPRLine = 2


class FontSize(enum.IntEnum):
    FontSizeNotDefined = -1
    FontSizeSmall = 0
    FontSizeMiddle = 1
    FontSizeBig = 2


class AnnotationFormat(enum.IntEnum):
    AnnotationFormatNotDefined = -1
    AnnotationFormatLeft = 0
    AnnotationFormatRight = 1
    AnnotationFormatCenter = 2


for _e in (LogLevel, PlaneType, Orientation, MPRType, RenderType, LayerType, FontSize, AnnotationFormat):
    globals().update(_e.__members__)


# ---- value types (core/Defines.h:36-80, core/Point.h, core/Direction.h) ----

class RGBA:
    def __init__(self, *args):
        if len(args) == 0:
            self.red = self.green = self.blue = self.alpha = 0.0
        elif len(args) == 3:
            self.red, self.green, self.blue = (float(a) for a in args)
            self.alpha = 1.0
        elif len(args) == 4:
            self.red, self.green, self.blue, self.alpha = (float(a) for a in args)
        else:
            raise TypeError("RGBA(), RGBA(r, g, b) or RGBA(r, g, b, a)")

    def Print(self) -> None:
        print(f"RGBA({self.red}, {self.green}, {self.blue}, {self.alpha})")

    def tuple(self):
        return (self.red, self.green, self.blue, self.alpha)


class Point2d:
    def __init__(self, x: float = 0.0, y: float = 0.0):
        self._c = [float(x), float(y)]

    def x(self):
        return self._c[0]

    def y(self):
        return self._c[1]

    def SetX(self, v):
        self._c[0] = float(v)

    def SetY(self, v):
        self._c[1] = float(v)


class Point3d:
    def __init__(self, x: float = 0.0, y: float = 0.0, z: float = 0.0):
        self._c = [float(x), float(y), float(z)]

    def x(self):
        return self._c[0]

    def y(self):
        return self._c[1]

    def z(self):
        return self._c[2]

    def SetX(self, v):
        self._c[0] = float(v)

    def SetY(self, v):
        self._c[1] = float(v)

    def SetZ(self, v):
        self._c[2] = float(v)

    def tuple(self):
        return tuple(self._c)


class Direction3d:
    """Direction3d() = (1, 0, 0); Direction3d(x, y, z) is normalised."""

    def __init__(self, *args):
        if len(args) == 0:
            self._c = (1.0, 0.0, 0.0)
        elif len(args) == 3:
            self._c = _mpr._dir(*(float(a) for a in args))
        else:
            raise TypeError("Direction3d() or Direction3d(x, y, z)")

    def x(self):
        return self._c[0]

    def y(self):
        return self._c[1]

    def z(self):
        return self._c[2]

    def tuple(self):
        return self._c


class BatchInfo:
    """core/BatchInfo.h:29-54 (the members without their m_ prefix); every
    field starts at zero, as the reference's memset constructor leaves it."""

    def __init__(self):
        self.dirH = (0.0, 0.0, 0.0)
        self.dirV = (0.0, 0.0, 0.0)
        self.ptCenter = (0.0, 0.0, 0.0)
        self.fLengthH = 0.0
        self.fLengthV = 0.0
        self.fPixelSpacing = 0.0
        self.fSliceThickness = 0.0
        self.fSliceDistance = 0.0
        self.nNum = 0
        self.mprType = MPRType.MPRTypeAverage

    def Width(self) -> int:
        return int(self.fLengthH / self.fPixelSpacing)

    def Height(self) -> int:
        return int(self.fLengthV / self.fPixelSpacing)

    def _series(self) -> _mpr.BatchInfo:
        def t(v):
            return v.tuple() if isinstance(v, (Direction3d, Point3d)) else tuple(float(c) for c in v)

        return _mpr.BatchInfo(t(self.dirH), t(self.dirV), t(self.ptCenter), self.fLengthH, self.fLengthV,
                              self.fPixelSpacing, self.fSliceThickness, self.fSliceDistance, self.nNum,
                              int(self.mprType))


class DeviceInfo:
    def GetCount(self, count: int = 0) -> bool:
        from . import _capi
        return _capi.load().mk_device_count() > 0


# ---- object state (core/ObjectInfo.h:19-43) ---------------------------------

class _ObjectInfo:
    def __init__(self):
        self.alpha, self.ww, self.wl = 1.0, 400.0, 40.0
        self.idx2rgba = {5: (0.8, 0.8, 0.8, 0.0), 95: (0.8, 0.8, 0.8, 0.8)}
        self.idx2alpha = {}


def _f32(x) -> float:
    return float(np.float32(x))


def _raw_memory(arr: np.ndarray) -> np.ndarray:
    """The array's elements in address order (what memcpy of the buffer
    pointer reads for a view with permuted strides)."""
    a = np.asarray(arr)
    order = sorted(range(a.ndim), key=lambda i: -abs(a.strides[i]))
    return np.ascontiguousarray(a.transpose(order)).reshape(-1)


def _rgba_map(ctrl) -> dict:
    out = {}
    for k, v in dict(ctrl).items():
        out[int(k)] = v.tuple() if isinstance(v, RGBA) else tuple(float(c) for c in v)
    return out


class HelloMonkey:
    """core/HelloMonkey.h:36-139 as pyMonkeyGL binds it."""

    def __init__(self):
        self._ctx = None
        self._log_level = LogLevel.LogLevelInfo
        self._dims = None                   # (W, H, D)
        self._vol = None                    # int16, x fastest
        self._has_mask = False              # the label mask lives on the device only
        self._spacing = (1.0, 1.0, 1.0)
        self._dir = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, -1.0))
        self._origin = (0.0, 0.0, 0.0)
        self._objects: dict = {}
        self._active = -1
        self._aw = [(0.0, 0.0, 0.0)] * MK_NUM_LABELS
        self._bg = RGBA()
        self._render_type = RenderType.RenderTypeVR
        self._ri = _P.RenderInfo()
        self._mpr = None
        self._cprinfo = _cpr.CPRInfo()
        self._cpr_color = RGBA(1.0, 1.0, 1.0)
        self._layers: dict = {}
        self._annotations: list = []
        self._plane_label = 0
        self._dirty_vol = self._dirty_lut = False
        self._encoders: dict = {}
        self._bufs: dict = {}
        self._series_renderer = None
        # the orientation figure: MarchingCube loads ./resources/human.stl at
        # construction (core/MarchingCube.cpp:297-298); without the file there
        # is no figure
        try:
            with open("./resources/human.stl", "rb") as f:
                self._orient_stl = f.read()
        except OSError:
            self._orient_stl = None
        self._orient_ws = None
        # annotation glyphs {(font size, byte): (h, w) uint8 mask}. The
        # reference compiles its character masks into the library
        # (AnnotationUtils::Init); here they are data: read from
        # ./resources/annotation_glyphs.npz when present (arrays named
        # "f<font>_c<byte>"), or registered with SetAnnotationGlyph
        self._glyphs: dict = {}
        self._glyphs_dirty = False
        try:
            with np.load("./resources/annotation_glyphs.npz", allow_pickle=False) as z:
                for name in z.files:
                    f, c = name[1:].split("_c")
                    self._glyphs[(int(f), int(c))] = np.ascontiguousarray(z[name], np.uint8)
            self._glyphs_dirty = True
        except (OSError, ValueError):
            pass

    # -- device sync ----------------------------------------------------------
    def _device(self):
        from ._capi import Context

        if self._ctx is None:
            self._ctx = Context(0)
            if self._orient_stl:
                try:
                    self._ctx.set_orientation_mesh(self._orient_stl)
                except ValueError:
                    self._orient_stl = None   # not a binary STL: no figure, as the reference logs and goes on
        c = self._ctx
        if self._glyphs_dirty:
            for (font, code), m in self._glyphs.items():
                c.set_glyph(font, code, m)
            self._glyphs_dirty = False
        if self._vol is None:
            raise RuntimeError("no volume: call SetVolumeArray first")
        W, H, D = self._dims
        if self._dirty_vol:
            c.set_volume(self._vol.reshape(D, H, W))   # drops the device mask, as SetVolumeData does
            self._dirty_vol = False
        if self._dirty_lut:
            # DataManager::UpdateTransferFunctions (core/DataManager.cpp:372-388)
            for label, info in self._objects.items():
                lut = _P.transfer_function(info.idx2rgba, info.idx2alpha)
                if lut is not None:
                    c.set_lut(label, lut)
            self._dirty_lut = False
        return c

    def _need_invert_z(self) -> bool:
        # VolumeInfo::Need2InvertZ (core/VolumeInfo.cpp:142-148)
        n = _cpr.cross(self._dir[0], self._dir[1])
        return _cpr.dot(n, self._dir[2]) > 0

    def _render_params(self, width=None, height=None):
        p = _P.render_params(self._dims, self._spacing, self._ri, {}, int(self._render_type),
                             invert_z=self._need_invert_z(), background=self._bg.tuple(), width=width,
                             height=height)
        for l in range(MK_NUM_LABELS):
            a, w, wl = self._aw[l]
            p.alpha_wwwl[l].alpha, p.alpha_wwwl[l].ww, p.alpha_wwwl[l].wl = a, w, wl
        return p

    def _encoder(self, w, h, ch):
        from .encode import Encoder

        key = (w, h, ch)
        if key not in self._encoders:
            self._encoders[key] = Encoder(w, h, ch, ctx=self._device())
        return self._encoders[key]

    def _dbuf(self, name: str, nbytes: int):
        """A device buffer of the context, kept per use and grown on demand.
        Every device call of this module runs on the context's stream and the
        results come back through synchronous copies, so a buffer is free
        again when the call that used it returns."""
        from ._capi import DeviceBuffer

        b = self._bufs.get(name)
        if b is None or b.nbytes < nbytes or b.ctx is not self._ctx:
            if b is not None:
                b.close()
            b = DeviceBuffer(self._ctx, max(int(nbytes), 1))
            self._bufs[name] = b
        return b

    # -- settings ---------------------------------------------------------------
    def SetLogLevel(self, level) -> None:
        self._log_level = LogLevel(int(level))

    def LoadVolumeFile(self, szFile: str) -> None:
        raise NotImplementedError("LoadVolumeFile reads through ITK, which this build does not include; "
                                  "pass the voxels with SetVolumeArray")

    def AddObjectMaskFile(self, szFile: str) -> int:
        raise NotImplementedError("AddObjectMaskFile reads through ITK, which this build does not include; "
                                  "pass the labels with AddNewObjectMaskArray")

    def SetVolumeArray(self, npData) -> bool:
        """DataManager::SetVolumeData (core/DataManager.cpp:144-165)."""
        a = np.asarray(npData)
        if a.ndim != 3 or min(a.shape) <= 0:
            return False
        self._active = -1
        self._objects = {}
        self._dims = tuple(int(s) for s in a.shape)
        self._vol = _raw_memory(a).astype(np.int16)
        self._has_mask = False
        self._mpr = _mpr.MPRInfo(self._dims, self._spacing)
        self._active = 0
        self._objects[0] = _ObjectInfo()
        self._dirty_vol = True
        return True

    def GetVolumeArray(self):
        W, H, D = self._dims
        return self._vol.reshape(W, H, D).copy()

    def AddNewObjectMaskArray(self, npData) -> int:
        """DataManager::AddNewObjectMask (core/DataManager.cpp:280-315); the
        merge runs on the device (mk_merge_object_mask, MK_MASK_ADD)."""
        a = _raw_memory(np.asarray(npData)).astype(np.uint8)
        if len(self._objects) <= 0:
            self._objects[0] = _ObjectInfo()
            label = 1
        elif len(self._objects) >= MAXOBJECTCOUNT + 1:
            return 0
        else:
            label = 1
            for k in sorted(self._objects)[1:]:
                if k != label:
                    break
                label += 1
        if self._vol is None or a.size < self._vol.size:
            return 0
        W, H, D = self._dims
        self._device().merge_mask(a[:self._vol.size].reshape(D, H, W), label, MK_MASK_ADD)
        self._has_mask = True
        self._objects[label] = copy.deepcopy(self._objects[self._active])
        self._active = label
        return label

    def UpdateMaskArray(self, npData, nLabel: int) -> bool:
        """DataManager::UpdateObjectMask (core/DataManager.cpp:324-335,
        VolumeInfo.cpp:118-133) as a device merge (MK_MASK_UPDATE)."""
        nLabel = int(nLabel)
        if nLabel not in self._objects or not self._has_mask:
            return False
        a = _raw_memory(np.asarray(npData)).astype(np.uint8)
        if a.size < self._vol.size:
            return False
        W, H, D = self._dims
        self._device().merge_mask(a[:self._vol.size].reshape(D, H, W), nLabel, MK_MASK_UPDATE)
        return True

    def GetMaskArray(self):
        """The merged label mask (W, H, D) view of an x-fastest buffer, like
        GetVolumeArray; None without a mask (DataManager::GetMaskData)."""
        if not self._has_mask:
            return None
        W, H, D = self._dims
        return self._device().get_mask().reshape(W, H, D)

    def SetSpacing(self, x: float, y: float, z: float) -> None:
        self._spacing = (float(x), float(y), float(z))
        self._cprinfo.set_spacing(self._spacing)
        if self._dims is not None:
            self._mpr = _mpr.MPRInfo(self._dims, self._spacing)

    def SetDirection(self, dirX, dirY, dirZ) -> None:
        self._dir = tuple(d.tuple() if isinstance(d, Direction3d) else _mpr._dir(*d) for d in (dirX, dirY, dirZ))

    def SetOrigin(self, pt) -> None:
        self._origin = pt.tuple() if isinstance(pt, Point3d) else tuple(float(c) for c in pt)

    def Reset(self) -> None:
        if self._dims is not None:
            self._mpr = _mpr.MPRInfo(self._dims, self._spacing)

    def SetColorBackground(self, clrBG) -> None:
        self._bg = clrBG

    # -- transfer functions, windows, alpha ---------------------------------------
    def SetTransferFunc(self, ctrlPoints, arg2=None, arg3=None) -> bool:
        """The four overloads of DataManager::SetControlPoints_TF
        (core/DataManager.cpp:337-370): (rgb), (rgb, label), (rgb, alpha),
        (rgb, alpha, label)."""
        if isinstance(arg2, dict):
            alpha, label = arg2, (self._active if arg3 is None else int(arg3))
        else:
            alpha, label = None, (self._active if arg2 is None else int(arg2))
        if label < 0 or label > MAXOBJECTCOUNT or label not in self._objects:
            return False
        self._objects[label].idx2rgba = _rgba_map(ctrlPoints)
        if alpha is not None:
            self._objects[label].idx2alpha = {int(k): float(v) for k, v in alpha.items()}
        self._dirty_lut = True
        return True

    def LoadTransferFunction(self, szFile: str) -> bool:
        """ObjectInfo::ReadFile (core/ObjectInfo.h:67-93) then SetVRWWWL and SetTransferFunc."""
        try:
            tok = open(szFile).read().split()
        except OSError:
            return False
        it = iter(tok)
        next(it); ww = float(next(it))
        next(it); wl = float(next(it))
        next(it); n = int(next(it))
        rgba = {}
        for _ in range(n):
            k = int(next(it))
            rgba[k] = tuple(_f32(next(it)) for _ in range(4))
        next(it); m = int(next(it))
        alpha = {}
        for _ in range(m):
            k = int(next(it))
            alpha[k] = _f32(next(it))
        return self.SetVRWWWL(ww, wl) and self.SetTransferFunc(rgba, alpha)

    def SaveTransferFunction(self, szFile: str) -> bool:
        """ObjectInfo::WriteFile (core/ObjectInfo.h:45-65) of the active object."""
        if self._active < 0 or not self._objects:
            return False
        info = self._objects[self._active]

        def g(v):
            return f"{_f32(v):g}"

        lines = [f"ww: {g(info.ww)}", f"wl: {g(info.wl)}", f"idx2rgba: {len(info.idx2rgba)}"]
        for k in sorted(info.idx2rgba):
            lines.append(f"{k} " + " ".join(g(c) for c in info.idx2rgba[k]) + " ")
        lines.append(f"idx2alpha: {len(info.idx2alpha)}")
        for k in sorted(info.idx2alpha):
            lines.append(f"{k} {g(info.idx2alpha[k])} ")
        with open(szFile, "w") as f:
            f.write("\n".join(lines) + "\n")
        return True

    def _update_alpha_wwwl(self) -> None:
        # DataManager::UpdateAlphaWWWL (core/DataManager.cpp:602-611)
        aw = list(self._aw)
        for label, info in self._objects.items():
            aw[label] = (_f32(info.alpha), _f32(info.ww), _f32(info.wl))
        self._aw = aw

    def SetVRWWWL(self, fWW: float, fWL: float, nLabel=None) -> bool:
        label = self._active if nLabel is None else int(nLabel)
        if label < 0 or label > MAXOBJECTCOUNT:
            return False
        if fWW <= 0:
            fWW = 1.0
        info = self._objects.setdefault(label, _ObjectInfo())
        info.ww, info.wl = _f32(fWW), _f32(fWL)
        self._update_alpha_wwwl()
        return True

    def SetObjectAlpha(self, fAlpha: float, nLabel=None) -> bool:
        label = self._active if nLabel is None else int(nLabel)
        if label < 0 or label > MAXOBJECTCOUNT:
            return False
        self._objects.setdefault(label, _ObjectInfo()).alpha = _f32(fAlpha)
        self._update_alpha_wwwl()
        return True

    # -- camera ---------------------------------------------------------------------
    def SetRenderType(self, type_) -> None:
        self._render_type = RenderType(int(type_))

    def Rotate(self, fxRotate: float, fyRotate: float) -> None:
        self._ri.rotate(fxRotate, fyRotate)

    def Zoom(self, ratio: float) -> float:
        return self._ri.zoom(ratio)

    def GetZoomRatio(self) -> float:
        return float(self._ri.total_scale)

    def Pan(self, fxShift: float, fyShift: float) -> None:
        self._ri.pan(fxShift, fyShift)

    def Anterior(self):
        self._ri.anterior()

    def Posterior(self):
        self._ri.posterior()

    def Left(self):
        self._ri.left()

    def Right(self):
        self._ri.right()

    def Head(self):
        self._ri.head()

    def Foot(self):
        self._ri.foot()

    def SetVRSize(self, nWidth: int, nHeight: int) -> bool:
        r = self._ri.set_vr_size(nWidth, nHeight)
        return True if r is None else bool(r)

    # -- MPR ---------------------------------------------------------------------------
    def _mpr_state(self):
        if self._mpr is None:
            raise RuntimeError("no volume: call SetVolumeArray first")
        return self._mpr

    def UpdateThickness(self, val: float) -> None:
        self._mpr_state().update_thickness(val)

    def SetThickness(self, val: float, planeType) -> None:
        self._mpr_state().set_thickness(val, int(planeType))

    def GetThickness(self, planeType) -> float:
        return self._mpr_state().get_thickness(int(planeType))

    def SetMPRType(self, type_) -> None:
        self._mpr_state().set_mpr_type(int(type_))

    def Browse(self, fDelta: float, planeType) -> None:
        self._mpr_state().browse(fDelta, int(planeType))

    def PanCrossHair(self, fx: float, fy: float, planeType) -> None:
        self._mpr_state().pan_crosshair(fx, fy, int(planeType))

    def RotateCrossHair(self, fAngle: float, planeType) -> None:
        self._mpr_state().rotate_crosshair(fAngle, int(planeType))

    def SetPlaneIndex(self, index: int, planeType) -> None:
        self._mpr_state().set_plane_index(int(index), int(planeType))

    def GetPlaneIndex(self, planeType) -> int:
        v = self._mpr_state().get_plane_index(int(planeType))
        return -1 if v is None else v

    def GetPlaneCurrentIndex(self, planeType) -> int:
        return self.GetPlaneIndex(planeType)

    def GetPlaneNumber(self, planeType) -> int:
        v = self._mpr_state().get_plane_number(int(planeType))
        return 0 if v is None else v

    def GetPlaneTotalNumber(self, planeType) -> int:
        return self.GetPlaneNumber(planeType)

    def GetCrossHairPoint(self, planeType) -> Point2d:
        v = self._mpr_state().get_crosshair_point(int(planeType))
        return Point2d(*(v if v is not None else (0.0, 0.0)))

    # -- CPR ---------------------------------------------------------------------------
    def SetCPRLinePatientArray(self, cprLineArray) -> bool:
        """DataManager::SetCPRLinePatient: points in patient mm through Patient2Voxel."""
        a = np.asarray(cprLineArray, np.float32)
        if a.ndim != 2 or a.shape[1] != 3:
            return False
        return self._cprinfo.set_cpr_line_patient(a.astype(np.float64), self._origin, *self._dir)

    def SetCPRLineVoxelArray(self, cprLineArray) -> bool:
        a = np.asarray(cprLineArray, np.float32)
        if a.ndim != 2 or a.shape[1] != 3:
            return False
        return self._cprinfo.set_cpr_line_voxel(a.astype(np.float64))

    def RotateCPR(self, angle: float, planeType) -> bool:
        p = int(planeType)
        which = {PlaneType.PlaneStretchedCPR: _cpr.PLANE_STRETCHED_CPR,
                 PlaneType.PlaneStraightenedCPR: _cpr.PLANE_STRAIGHTENED_CPR}.get(p)
        return self._cprinfo.rotate_cpr(angle, which) if which else False

    def SetCPRLineColor(self, clr) -> None:
        self._cpr_color = clr

    # -- layers and annotations ------------------------------------------------------------
    def ShowPlaneInVR(self, bShow: bool) -> None:
        """DataManager::ShowPlaneInVR (core/DataManager.cpp:769-812): three slabs
        added as a label at 0.6 alpha."""
        if not bShow or self._dims is None:
            return
        W, H, D = self._dims
        m = np.zeros((D, H, W), np.uint8)
        m[200:203] = 1
        m[:, 200:203, :] = 1
        m[:, :, 250:253] = 1
        self._plane_label = self.AddNewObjectMaskArray(m.reshape(-1))
        self.SetVRWWWL(1000, -40)
        self.SetObjectAlpha(0.6, self._plane_label)

    def AddAnnotation(self, planeType, txt, x, y, fontSize, annoFormat, clr) -> bool:
        self._annotations.append((PlaneType.PlaneVR, txt, x, y, fontSize, annoFormat, clr))
        return True

    def SetAnnotationGlyph(self, fontSize, char, mask) -> None:
        """The mask of one character (a one-byte string or a byte value) in one
        font size, (height, width) uint8; None removes it. No reference
        counterpart: the reference's masks are compiled into its library."""
        code = char if isinstance(char, int) else ord(char.encode("latin-1") if isinstance(char, str) else char)
        key = (int(fontSize), int(code))
        if mask is None:
            self._glyphs.pop(key, None)
        else:
            self._glyphs[key] = np.ascontiguousarray(mask, np.uint8)
        if self._ctx is not None:
            self._ctx.set_glyph(key[0], key[1], None if mask is None else self._glyphs[key])
        else:
            self._glyphs_dirty = True   # registered with the context when it is made

    def RemovePlaneAnnotations(self, planeType) -> bool:
        self._annotations = [a for a in self._annotations if a[0] != int(planeType)]
        return True

    def RemoveAllAnnotations(self) -> bool:
        self._annotations = []
        return True

    def EnableLayer(self, layerType, bEnable: bool) -> bool:
        self._layers[int(layerType)] = bool(bEnable)
        return True

    def _image_layer(self) -> bool:
        return self._layers.get(int(LayerType.LayerTypeImage), True)

    # -- output ------------------------------------------------------------------------
    def _vr_device(self):
        """The VR frame at the VR size in a device buffer (cu_render's path)."""
        from ._capi import DeviceBuffer

        c = self._device()
        w, h = self._ri.width, self._ri.height
        p = self._render_params(w, h)
        out = self._dbuf("vr", w * h * 3)
        st = c.stream
        c.render_async(p, out, stream=st)
        if self._orient_stl:
            # VRImageLayer::GetRGBData ends with MergeOrientationBox (core/VRImageLayer.cpp:99)
            ws = self._dbuf("orient", c.orientation_workspace_bytes())
            c.merge_orientation_box_async(self._ri.rot, out, w, h, ws, stream=st)
        # AnnotationLayer::GetRGBData (core/AnnotationLayer.cpp:41-61): the VR
        # plane's annotations in order, on frames wider and taller than 10
        if self._layers.get(int(LayerType.LayerTypeAnnotation), True) and w > 10 and h > 10:
            for plane, txt, x, y, fs, fmt, clr in self._annotations:
                if plane == PlaneType.PlaneVR:
                    c.draw_text_async(str(txt).encode("utf-8"), int(x), int(y), int(fs), int(fmt),
                                      clr.tuple()[:3], out, w, h, stream=st)
        # the CPR line layer draws the CPR line's voxel points projected into the image
        line = self._cprinfo.line
        if self._layers.get(int(LayerType.LayerTypeCPRLine), True) and len(line) >= 2:
            xy = DeviceBuffer.of(c, self._voxel_to_vr(line, w, h))
            c.draw_polyline_async(xy, len(line), self._cpr_color.tuple()[:3], out, w, h, stream=st)
            out._keep = xy
        return out, w, h

    def _voxel_to_vr(self, pts, w: int, h: int) -> np.ndarray:
        """DataManager::TransferVoxel2ImageInVR (core/DataManager.cpp:236-258)
        for each voxel point: (x, y) float pairs in the w x h VR image."""
        W, H, D = self._dims
        sx, sy, sz = self._spacing
        max_len = float(np.float32(max(W * sx, max(H * sy, D * sz))))
        maxper = (max_len / (W * sx), max_len / (H * sy), max_len / (D * sz))
        mt = [float(v) for v in np.asarray(self._ri.m_t, np.float32).reshape(9)]
        inv = self._need_invert_z()
        out = np.empty((len(pts), 2), np.float32)
        for k, p in enumerate(pts):
            x, y, z = p[0] / W, p[1] / H, p[2] / D
            if inv:
                z = 1.0 - z
            x, y, z = (x - 0.5) / maxper[0], (y - 0.5) / maxper[1], (z - 0.5) / maxper[2]
            rx = mt[0] * x + mt[1] * y + mt[2] * z
            rz = mt[6] * x + mt[7] * y + mt[8] * z
            out[k, 0] = (rx + 0.5) * w + float(self._ri.x_translate)
            out[k, 1] = (rz + 0.5) * h + float(self._ri.y_translate)
        return out

    def GetVRArray(self, nWidth: int, nHeight: int):
        if not self._image_layer():
            raise RuntimeError("the image layer is disabled")
        w, h = self._ri.width, self._ri.height
        if nWidth * nHeight != w * h:
            raise ValueError(f"GetVRArray reads the VR frame ({w}x{h}); ask for a shape of the same size")
        out, w, h = self._vr_device()
        return out.to_host(np.uint8, w * h * 3).reshape(3, nWidth, nHeight)

    def _vr_png(self) -> bytes:
        out, w, h = self._vr_device()
        enc = self._encoder(w, h, 3)
        enc.encode_async(out, w, h, 3, stream=self._ctx.stream, base64=False)
        return enc.png_bytes()

    def GetVRData_pngString(self, *ignored) -> str:
        if not self._image_layer():
            return ""
        out, w, h = self._vr_device()
        enc = self._encoder(w, h, 3)
        enc.encode_async(out, w, h, 3, stream=self._ctx.stream)
        return enc.base64_text()

    def GetVRData_png(self, nWidth: int = 0, nHeight: int = 0) -> list:
        if not self._image_layer():
            return []
        return list(self._vr_png())

    def SaveVR2Png(self, szFile: str, nWidth: int = 0, nHeight: int = 0) -> None:
        with open(szFile, "wb") as f:
            f.write(self._vr_png())

    def _plane_device(self, planeType):
        """(int16 device image, width, height) of an MPR or CPR plane, or None."""
        from ._capi import DeviceBuffer

        p = int(planeType)
        c = self._device()
        st = c.stream
        if PlaneType.PlaneAxial <= p <= PlaneType.PlaneCoronalOblique:
            q = self._mpr_state().plane_params(p, invert_z=self._need_invert_z(), sampling=MK_SAMPLE_LINEAR)
            out = self._dbuf("plane", q.width * q.height * 2)
            c.render_plane_async(q, out, stream=st)
            return out, q.width, q.height
        if p in (PlaneType.PlaneStretchedCPR, PlaneType.PlaneStraightenedCPR):
            which = _cpr.PLANE_STRETCHED_CPR if p == PlaneType.PlaneStretchedCPR else _cpr.PLANE_STRAIGHTENED_CPR
            g = self._cprinfo.get_cpr_info(which)
            if g is None:
                return None
            pts, dirs, w, h = g
            dp = DeviceBuffer.of(c, np.asarray(pts, np.float64))
            dd = DeviceBuffer.of(c, np.asarray(dirs, np.float64))
            out = self._dbuf("plane", w * h * 2)
            c.render_cpr_async(w, h, dp, dd, self._need_invert_z(), MK_SAMPLE_LINEAR, out, stream=st)
            out._keep = (dp, dd)
            return out, w, h
        return None

    def GetPlaneData(self, planeType):
        """(int16 image (h, w), w, h) as Render::GetPlaneData gives it, or None."""
        r = self._plane_device(planeType)
        if r is None:
            return None
        out, w, h = r
        return out.to_host(np.int16, w * h).reshape(h, w), w, h

    def _int16_png_b64(self, dev, w, h) -> str:
        enc = self._encoder(w // 2, h, 4)
        enc.encode_async(dev, w // 2, h, 4, stream=self._ctx.stream)
        return enc.base64_text()

    def GetPlaneData_pngString(self, planeType) -> str:
        if not self._image_layer():
            return ""
        r = self._plane_device(planeType)
        if r is None:
            return ""
        out, w, h = r
        return self._int16_png_b64(out, w, h)

    def GetOriginData_pngString(self, slice_: int) -> str:
        """An original axial slice, -2048 outside the mask when there is one
        (core/HelloMonkey.cpp:261-319)."""
        W, H, D = self._dims
        c = self._device()
        dev = self._dbuf("origin", W * H * 2)
        c.origin_slice_async(int(slice_), dev, stream=c.stream)
        return self._int16_png_b64(dev, W, H)

    # -- batch series (core/HelloMonkey.cpp:415-420, 614-619) ---------------------
    def GetBatchDirection3D(self, fAngle: float, planeType):
        """(dirH, dirV) of a batch plane: the plane's dirH turned by fAngle
        degrees about its normal, and the normal (DataManager::GetBatchDirection3D);
        None for a plane without MPR state."""
        d = self._mpr_state().get_direction_3d(int(planeType))
        if d is None:
            return None
        h, v = _mpr.batch_direction_3d(d[0], d[1], float(fAngle))
        return Direction3d(*h), Direction3d(*v)

    def GetBatchData(self, batchInfo) -> list:
        """The images of a batch series as (height, width) int16 arrays, one
        per slice, or [] for an empty or degenerate series. The reference
        leaves this unimplemented (core/Render.cpp:117-120); the series is
        nNum parallel slabs centred on ptCenter, fSliceDistance apart along
        dirH x dirV, rendered in batches on the GPU."""
        from .series import SeriesRenderer

        params = _mpr.batch_plane_params(batchInfo._series(), self._spacing, invert_z=self._need_invert_z(),
                                         sampling=MK_SAMPLE_LINEAR)
        if not params:
            return []
        c = self._device()
        if self._series_renderer is None:
            self._series_renderer = SeriesRenderer(c)
        arr = self._series_renderer.render(params)
        return [arr[k] for k in range(arr.shape[0])]

    def Transfer2Base64Array(self, npData) -> None:
        """Prints the base64 of width * height bytes as a quoted line (core/HelloMonkey.cpp:54-61)."""
        a = np.asarray(npData)
        n = a.shape[0] * a.shape[1]
        print(f'"{_b64.b64encode(_raw_memory(a).astype(np.uint8)[:n].tobytes()).decode()}",')


__all__ = ["HelloMonkey", "RGBA", "Point2d", "Point3d", "Direction3d", "DeviceInfo", "BatchInfo", "LogLevel",
           "PlaneType", "Orientation", "MPRType", "RenderType", "LayerType", "FontSize", "AnnotationFormat"]
__all__ += [m for e in (LogLevel, PlaneType, Orientation, MPRType, RenderType, LayerType, FontSize,
                        AnnotationFormat) for m in e.__members__]
