"""The pyMonkeyGL drop-in needs no torch.

A fresh Python process with torch made unimportable (sys.modules["torch"] =
None) drives the module the way the reference's examples drive pyMonkeyGL
(examples/examples.py): volume, transfer function, window, a label mask,
rotation, then the VR frame as a base64 PNG and as an array, MPR planes, a
straightened CPR, an original slice and a batch series. Device memory comes
from the library (mk_device_alloc) and results from its synchronous copies.
Each image must equal the one-request entry points' image bit for bit, and
the VR frame the FP64 oracle within 1/255 on >= 99.9% of pixels; the process
must end without any torch module loaded.
"""
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent

SCRIPT = r'''
import base64, sys
sys.modules["torch"] = None          # any import of torch now raises ImportError
sys.path.insert(0, ROOT)
import numpy as np
from monkeygl_amd import pyMonkeyGL as mk
from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_SAMPLE_LINEAR
from oracle import oracle as O
from tests.png_util import read_png

spec = W.VolumeSpec("notorch", (96, 80, 72), (0.8, 0.8, 1.0), 6, (3, 12), (300, 1800), n_vessels=3, seed=41)
vol = W.make_volume(spec)
hm = mk.HelloMonkey()
hm.SetVolumeArray(vol.swapaxes(2, 0))
hm.SetSpacing(*spec.spacing)
hm.SetTransferFunc({5: mk.RGBA(0.8, 0, 0, 0), 90: mk.RGBA(0.8, 0.8, 0.8, 0.8)})
hm.SetVRWWWL(500, 100)
m = W.ellipsoid_mask(spec.dims, 1, seed=3)
assert hm.AddNewObjectMaskArray((m == 1).astype(np.uint8).swapaxes(2, 0)) == 1
hm.SetVRSize(200, 160)
hm.Rotate(30, 20)

# VR: base64 PNG, array, against the plain entry point and the oracle
img, info = read_png(base64.b64decode(hm.GetVRData_pngString()))
p = hm._render_params(hm._ri.width, hm._ri.height)
assert np.array_equal(img, hm._ctx.render(p)), "VR png != mk_render"
luts = {l: P.transfer_function(o.idx2rgba, o.idx2alpha) for l, o in hm._objects.items()}
luts = {l: v for l, v in luts.items() if v is not None}
ref = O.render(p, vol, luts, mask_zyx=m)
d = np.abs(img.astype(int) - ref.astype(int)).max(-1)
assert (d <= 1).mean() >= 0.999, (d <= 1).mean()
arr = hm.GetVRArray(200, 160)
assert np.array_equal(arr.reshape(-1), img.reshape(-1)), "GetVRArray"

# MPR planes and their PNG strings
for plane in (mk.PlaneAxial, mk.PlaneSagittal, mk.PlaneCoronal):
    got, w, h = hm.GetPlaneData(plane)
    q = hm._mpr_state().plane_params(int(plane), invert_z=hm._need_invert_z(), sampling=MK_SAMPLE_LINEAR)
    assert np.array_equal(got, hm._ctx.render_plane(q)), ("plane", plane)
    rgba, _ = read_png(base64.b64decode(hm.GetPlaneData_pngString(plane)))
    assert np.array_equal(rgba.reshape(-1).view(np.int16), got.reshape(-1)), ("plane png", plane)

# stretched and straightened CPR along a line through the volume
from monkeygl_amd import cpr as G
line = np.stack([np.linspace(20, 70, 60), np.full(60, 40.0), np.linspace(10, 60, 60)], 1).astype(np.float32)
assert hm.SetCPRLineVoxelArray(line)
for plane, which in ((mk.PlaneStretchedCPR, G.PLANE_STRETCHED_CPR), (mk.PlaneStraightenedCPR, G.PLANE_STRAIGHTENED_CPR)):
    got, w, h = hm.GetPlaneData(plane)
    pts, dirs, w2, h2 = hm._cprinfo.get_cpr_info(which)
    assert (w, h) == (w2, h2)
    assert np.array_equal(got, hm._ctx.render_cpr(w, pts, dirs, hm._need_invert_z(), MK_SAMPLE_LINEAR)), plane

# an original slice and a batch series
o_img, _ = read_png(base64.b64decode(hm.GetOriginData_pngString(30)))
assert o_img.size > 0
from monkeygl_amd import mpr as M
dh, dv = hm.GetBatchDirection3D(25.0, mk.PlaneSagittal)
bi = mk.BatchInfo()
bi.dirH, bi.dirV = dh, dv
bi.ptCenter = mk.Point3d(38.4, 36.0, 43.2)
bi.fLengthH, bi.fLengthV, bi.fPixelSpacing = 64.0, 48.0, 0.8
bi.fSliceThickness, bi.fSliceDistance, bi.nNum = 2.4, 2.0, 6
bi.mprType = mk.MPRTypeMIP
imgs = hm.GetBatchData(bi)
params = M.batch_plane_params(bi._series(), spec.spacing, invert_z=hm._need_invert_z())
assert len(imgs) == 6
for k, q in enumerate(params):
    assert np.array_equal(imgs[k], hm._ctx.render_plane(q)), ("series", k)

assert sys.modules.get("torch") is None
assert not [k for k in sys.modules if k.startswith("torch.")], "a torch submodule was loaded"
print("NOTORCH_OK", img.shape, float((d <= 1).mean()))
'''


def test_pymonkeygl_runs_without_torch():
    script = "ROOT = " + repr(str(ROOT)) + "\n" + SCRIPT
    run = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=300, cwd=str(ROOT))
    print(run.stdout[-2000:], run.stderr[-3000:])
    assert run.returncode == 0 and "NOTORCH_OK" in run.stdout
