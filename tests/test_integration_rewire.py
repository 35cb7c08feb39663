"""The reference-side rewiring compiles, links and runs (no GPU needed).

tests/integration/rewire.cpp holds the call sites a maintainer rewires in
MonkeyGL onto the C ABI (CuDataManager's constructor, destructor,
CopyVolumeData, CopyMaskData and SetTransferFunction; VRImageLayer's
parameter fill and mk_render; MPRImageLayer's mk_render_plane;
CPRImageLayer's mk_render_cpr). It is compiled against the reference's own
core/Defines.h, so the structs and enums handed across are the reference's
(static_asserts check their layouts and values against include/monkeygl_hip.h),
and linked against libmonkeygl_hip.so. Without a GPU, creating the context
must return MK_ERR_HIP with a message and the program must carry on and exit
normally; with one, every call site renders.

The reference's headers are read where they are mounted; elsewhere the test
is skipped.
"""
import os
import subprocess
from pathlib import Path

import pytest

from monkeygl_amd import _capi

ROOT = Path(__file__).resolve().parent.parent
REF_CORE = Path(os.environ.get("MONKEYGL_REFERENCE", "/root/reference")) / "core"


@pytest.mark.skipif(not (REF_CORE / "Defines.h").exists(), reason="the reference's headers are not mounted here")
def test_rewired_call_sites_compile_link_and_run(tmp_path):
    lib = _capi.load()   # builds nothing; fails if the library is missing
    exe = tmp_path / "rewire"
    cmd = ["g++", "-std=c++11", "-I", str(REF_CORE), "-I", str(ROOT / "include"),
           str(ROOT / "tests" / "integration" / "rewire.cpp"), "-L", str(ROOT / "monkeygl_amd"), "-lmonkeygl_hip",
           f"-Wl,-rpath,{ROOT / 'monkeygl_amd'}", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(run.stdout)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
    if lib.mk_device_count() == 0:
        assert f"status {_capi.MK_ERR_HIP}:" in run.stdout and "no HIP device" in run.stdout
    else:
        assert run.stdout.startswith("vr ") and " mpr " in run.stdout and " cpr " in run.stdout
