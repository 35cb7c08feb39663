"""Kernel-argument producers against golden vectors from the reference.

tests/golden/*.json were written by running the reference's own
TransferFunction.cpp, Methods.cpp and RenderInfo.cpp (compiled unmodified by
oracle/build_ref.sh); the restatements in monkeygl_amd.params must reproduce
them bit for bit (float32 equality).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from monkeygl_amd import params as P

GOLDEN = Path(__file__).resolve().parent / "golden"


def _tf_cases():
    return json.loads((GOLDEN / "tf_luts.json").read_text())


@pytest.mark.parametrize("case", _tf_cases(), ids=lambda c: c["name"])
def test_transfer_function_matches_reference(case):
    pts = {int(k): tuple(v) for k, v in case["pts"].items()}
    alpha = {int(k): v for k, v in case["alpha"].items()} or None
    got = P.transfer_function(pts, alpha)
    if case["lut"] is None:
        assert got is None
    else:
        ref = np.array(case["lut"], np.float32)
        assert got is not None and got.shape == ref.shape
        assert np.array_equal(got, ref), np.abs(got - ref).max()


def test_render_info_sequences_match_reference():
    data = json.loads((GOLDEN / "render_info.json").read_text())
    for seq in data["sequences"]:
        info = P.RenderInfo()
        for op, st in zip(seq["ops"], seq["states"]):
            kind = op[0]
            if kind == "rotate":
                info.rotate(op[1], op[2])
            elif kind == "zoom":
                info.zoom(op[1])
            elif kind == "pan":
                info.pan(op[1], op[2])
            elif kind == "preset":
                [info.anterior, info.posterior, info.left, info.right, info.head, info.foot][op[1]]()
            for key, arr in (("m", info.m), ("m_t", info.m_t), ("rot", info.rot), ("rot_t", info.rot_t)):
                assert np.array_equal(arr, np.array(st[key], np.float32)), (seq["name"], op, key)
            assert np.float32(info.x_translate) == np.float32(st["x_translate"])
            assert np.float32(info.y_translate) == np.float32(st["y_translate"])
            assert np.float32(info.total_scale) == np.float32(st["scale"])


def test_vr_size_rule_matches_reference():
    data = json.loads((GOLDEN / "render_info.json").read_text())
    for w, h, ow, oh in data["vr_sizes"]:
        info = P.RenderInfo()
        info.set_vr_size(w, h)
        assert (info.width, info.height) == (ow, oh), (w, h)


def test_transform_is_a_scaled_rotation():
    """Closed form: after any rotation sequence at zoom s, M = R / s with R
    orthonormal, and M_t is its inverse (M_t . M = I)."""
    info = P.RenderInfo()
    for a, b in [(20, 35), (-75, 110), (5.5, -3.25)]:
        info.rotate(a, b)
    info.zoom(1.6)
    M = info.m.reshape(3, 3).astype(np.float64)
    Mt = info.m_t.reshape(3, 3).astype(np.float64)
    s = 1.0 / 1.6
    assert np.allclose(M @ M.T, s * s * np.eye(3), atol=1e-5)
    assert np.allclose(Mt @ M, np.eye(3), atol=1e-5)


def test_init_common_closed_form():
    sp, sv, mls = P.init_common((0.7422, 0.7422, 1.0), (512, 512, 1559))
    assert np.isclose(sv[2], 1 / 1559, rtol=1e-7)
    # the largest physical extent has max_len_spacing 1
    assert mls[2] == np.float32(1.0)
    assert np.isclose(mls[0], 1559.0 / (512 * 0.7422), rtol=1e-6)


def test_need_to_invert_z():
    assert not P.need_to_invert_z()                       # reference default directions
    assert P.need_to_invert_z((1, 0, 0), (0, 1, 0), (0, 0, 1))


def test_reference_library_agrees_when_present():
    """Where the reference is mounted and oracle/build_ref.sh has run, also
    compare against the compiled reference directly on fresh random inputs."""
    lib = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libmkref.so"
    if not lib.exists():
        pytest.skip("oracle/_ref not built")
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    L = mg.load()
    rng = np.random.default_rng(99)
    for _ in range(50):
        n = int(rng.integers(2, 10))
        pos = rng.choice(np.arange(-5, 105), size=n, replace=False)
        pts = {int(p): tuple(float(x) for x in rng.uniform(0, 1, 4).astype(np.float32)) for p in pos}
        ref = mg.ref_tf(L, pts)
        got = P.transfer_function(pts)
        assert (ref is None) == (got is None)
        if ref is not None:
            assert np.array_equal(got, ref)
