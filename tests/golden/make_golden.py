"""Writes golden vectors for the kernel-argument producers.

Runs the reference's own TransferFunction.cpp, Methods.cpp and RenderInfo.cpp,
compiled unmodified by oracle/build_ref.sh into oracle/_ref/libmkref.so, on a
fixed set of inputs and stores inputs and outputs as JSON next to this script:

  tf_luts.json        control points (+ optional alpha points) -> 100-entry LUT
  render_info.json    RenderInfo operation sequences -> matrices after each op
  cpr_geometry.json   CPR geometry helpers (direction, cross, RotatePoint3D,
                      Projection_Point2Plane, Distance_Point2Line,
                      PatientDistanceTo) on seeded inputs
  mpr_geometry.json   MPR plane helpers ((dirH, dirV) overloads of
                      Distance_Point2Plane / Projection_Point2Plane,
                      GetTransferPoint, PlaneInfo normal and left-top corner)

Usage: sh oracle/build_ref.sh && python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_LIB = HERE.parent.parent / "oracle" / "_ref" / "libmkref.so"


def load():
    L = C.CDLL(str(REF_LIB))
    vp, i, f = C.c_void_p, C.c_int, C.c_float
    L.ref_transfer_function.restype = i
    L.ref_transfer_function.argtypes = [vp, vp, i, vp, vp, i, vp, vp]
    L.ref_render_info_new.restype = vp
    L.ref_render_info_free.argtypes = [vp]
    L.ref_render_info_rotate.argtypes = [vp, f, f]
    L.ref_render_info_zoom.restype = f
    L.ref_render_info_zoom.argtypes = [vp, f]
    L.ref_render_info_pan.argtypes = [vp, f, f]
    L.ref_render_info_preset.argtypes = [vp, i]
    L.ref_render_info_set_vr_size.argtypes = [vp, i, i, vp, vp]
    L.ref_render_info_get.argtypes = [vp, vp, vp, vp, vp, vp]
    L.ref_direction.argtypes = [vp, vp, vp]
    L.ref_cross.argtypes = [vp, vp, vp, vp, vp]
    L.ref_rotate_point.argtypes = [vp, vp, vp, vp, f, vp]
    L.ref_projection_point_to_plane.argtypes = [vp, vp, vp, vp, vp]
    L.ref_distance_point_to_line.restype = C.c_double
    L.ref_distance_point_to_line.argtypes = [vp, vp, vp, vp]
    L.ref_patient_distance.restype = C.c_double
    L.ref_patient_distance.argtypes = [vp, vp, vp]
    L.ref_distance_point_to_plane_hv.restype = C.c_double
    L.ref_distance_point_to_plane_hv.argtypes = [vp] * 6
    L.ref_projection_point_to_plane_hv.argtypes = [vp] * 7
    L.ref_transfer_point.argtypes = [vp, vp, vp]
    L.ref_plane_info.argtypes = [vp, vp, vp, vp, i, i, C.c_double, vp, vp, vp]
    return L


def mpr_geometry_cases(L):
    rng = np.random.default_rng(78)

    def v():
        return rng.uniform(-50.0, 550.0, 3)

    def hx(a):
        return [float(x).hex() for x in np.atleast_1d(a)]

    out = {"dist_plane_hv": [], "project_hv": [], "transfer": [], "plane_info": []}
    o, o2 = np.zeros(3), np.zeros(3)
    for _ in range(48):
        pt, h0, h1, v0, v1, pp = (v() for _ in range(6))
        r = L.ref_distance_point_to_plane_hv(pt.ctypes.data, h0.ctypes.data, h1.ctypes.data, v0.ctypes.data,
                                             v1.ctypes.data, pp.ctypes.data)
        out["dist_plane_hv"].append({"pt": hx(pt), "h0": hx(h0), "h1": hx(h1), "v0": hx(v0), "v1": hx(v1),
                                     "pp": hx(pp), "out": hx(r)})
        L.ref_projection_point_to_plane_hv(pt.ctypes.data, h0.ctypes.data, h1.ctypes.data, v0.ctypes.data,
                                           v1.ctypes.data, pp.ctypes.data, o.ctypes.data)
        out["project_hv"].append({"pt": hx(pt), "h0": hx(h0), "h1": hx(h1), "v0": hx(v0), "v1": hx(v1),
                                  "pp": hx(pp), "out": hx(o)})
        m = rng.uniform(-1.0, 1.0, 9)
        L.ref_transfer_point(m.ctypes.data, pt.ctypes.data, o.ctypes.data)
        out["transfer"].append({"m": hx(m), "pt": hx(pt), "out": hx(o)})
        w, h = int(rng.integers(1, 1200)), int(rng.integers(1, 1200))
        ps = float(rng.uniform(0.2, 1.5))
        c = v()
        L.ref_plane_info(h0.ctypes.data, h1.ctypes.data, v0.ctypes.data, v1.ctypes.data, w, h, ps, c.ctypes.data,
                         o.ctypes.data, o2.ctypes.data)
        out["plane_info"].append({"h0": hx(h0), "h1": hx(h1), "v0": hx(v0), "v1": hx(v1), "w": w, "h": h,
                                  "ps": float(ps).hex(), "center": hx(c), "norm": hx(o), "left_top": hx(o2)})
    return out


def cpr_geometry_cases(L):
    """Seeded points in voxel-like ranges; each helper's output (doubles, as
    hex strings so the JSON round trip is exact)."""
    rng = np.random.default_rng(77)

    def v():
        return rng.uniform(-50.0, 550.0, 3)

    def hx(a):
        return [float(x).hex() for x in np.atleast_1d(a)]

    out = {"direction": [], "cross": [], "rotate": [], "project": [], "dist_line": [], "patient_dist": []}
    o = np.zeros(3)
    for _ in range(64):
        a, b = v(), v()
        L.ref_direction(a.ctypes.data, b.ctypes.data, o.ctypes.data)
        out["direction"].append({"a": hx(a), "b": hx(b), "out": hx(o)})
        a0, a1, b0, b1 = v(), v(), v(), v()
        L.ref_cross(a0.ctypes.data, a1.ctypes.data, b0.ctypes.data, b1.ctypes.data, o.ctypes.data)
        out["cross"].append({"a0": hx(a0), "a1": hx(a1), "b0": hx(b0), "b1": hx(b1), "out": hx(o)})
        pt, x0, x1, c = v(), v(), v(), v()
        theta = float(rng.uniform(-7.0, 7.0))
        L.ref_rotate_point(pt.ctypes.data, x0.ctypes.data, x1.ctypes.data, c.ctypes.data, theta, o.ctypes.data)
        out["rotate"].append({"pt": hx(pt), "ax0": hx(x0), "ax1": hx(x1), "center": hx(c),
                              "theta": float(np.float32(theta)).hex(), "out": hx(o)})
        pt, n0, n1, pp = v(), v(), v(), v()
        L.ref_projection_point_to_plane(pt.ctypes.data, n0.ctypes.data, n1.ctypes.data, pp.ctypes.data,
                                        o.ctypes.data)
        out["project"].append({"pt": hx(pt), "n0": hx(n0), "n1": hx(n1), "pp": hx(pp), "out": hx(o)})
        pt, d0, d1, pl = v(), v(), v(), v()
        r = L.ref_distance_point_to_line(pt.ctypes.data, d0.ctypes.data, d1.ctypes.data, pl.ctypes.data)
        out["dist_line"].append({"pt": hx(pt), "d0": hx(d0), "d1": hx(d1), "pl": hx(pl), "out": hx(r)})
        a, b, sp = v(), v(), rng.uniform(0.3, 1.5, 3)
        r = L.ref_patient_distance(a.ctypes.data, b.ctypes.data, sp.ctypes.data)
        out["patient_dist"].append({"a": hx(a), "b": hx(b), "sp": hx(sp), "out": hx(r)})
    return out


def ref_tf(L, pts: dict, alpha: dict | None = None):
    pos = np.array(sorted(pts), np.int32)
    rgba = np.array([pts[k] for k in sorted(pts)], np.float32).reshape(-1)
    alpha = alpha or {}
    apos = np.array(sorted(alpha), np.int32) if alpha else np.zeros(1, np.int32)
    aval = np.array([alpha[k] for k in sorted(alpha)], np.float32) if alpha else np.zeros(1, np.float32)
    out = np.zeros(4 * 1024, np.float32)
    n = C.c_int(0)
    ok = L.ref_transfer_function(pos.ctypes.data, rgba.ctypes.data, len(pos), apos.ctypes.data, aval.ctypes.data,
                                 len(alpha), out.ctypes.data, C.byref(n))
    return out[: 4 * n.value].reshape(-1, 4) if ok else None


def ref_render_info_run(L, ops):
    h = L.ref_render_info_new()
    states = []
    for op in ops:
        kind = op[0]
        if kind == "rotate":
            L.ref_render_info_rotate(h, op[1], op[2])
        elif kind == "zoom":
            L.ref_render_info_zoom(h, op[1])
        elif kind == "pan":
            L.ref_render_info_pan(h, op[1], op[2])
        elif kind == "preset":
            L.ref_render_info_preset(h, op[1])
        elif kind == "size":
            ow, oh = C.c_int(0), C.c_int(0)
            L.ref_render_info_set_vr_size(h, op[1], op[2], C.byref(ow), C.byref(oh))
        bufs = [np.zeros(9, np.float32) for _ in range(4)]
        xyz = np.zeros(3, np.float32)
        L.ref_render_info_get(h, *[b.ctypes.data for b in bufs], xyz.ctypes.data)
        states.append({"m": bufs[0].tolist(), "m_t": bufs[1].tolist(), "rot": bufs[2].tolist(),
                       "rot_t": bufs[3].tolist(), "x_translate": float(xyz[0]), "y_translate": float(xyz[1]),
                       "scale": float(xyz[2])})
    L.ref_render_info_free(h)
    return states


def ref_vr_size(L, w, h):
    r = L.ref_render_info_new()
    ow, oh = C.c_int(0), C.c_int(0)
    L.ref_render_info_set_vr_size(r, w, h, C.byref(ow), C.byref(oh))
    L.ref_render_info_free(r)
    return ow.value, oh.value


def tf_cases():
    rng = np.random.default_rng(1234)
    cases = [
        {"name": "object_default", "pts": {5: (0.8, 0.8, 0.8, 0.0), 95: (0.8, 0.8, 0.8, 0.8)}},
        # trfns/coro.txt control points
        {"name": "coro", "pts": {5: (0.8, 0.0, 0.0, 0.0), 90: (0.8, 0.8, 0.8, 0.8)}},
        {"name": "ends_exact", "pts": {0: (0.0, 0.1, 0.2, 0.0), 99: (1.0, 0.9, 0.8, 1.0)}},
        {"name": "one_point", "pts": {40: (0.5, 0.5, 0.5, 0.5)}},
        {"name": "all_out_of_range", "pts": {-5: (1, 1, 1, 1), 120: (0, 0, 0, 0)}},
        {"name": "partly_out_of_range", "pts": {-3: (1, 0, 0, 1), 30: (0, 1, 0, 0.3), 70: (0, 0, 1, 0.6), 140: (1, 1, 1, 1)}},
        {"name": "adjacent", "pts": {10: (1, 0, 0, 0.1), 11: (0, 1, 0, 0.9), 12: (0, 0, 1, 0.2)}},
        {"name": "alpha_map", "pts": {0: (0.9, 0.2, 0.1, 1.0), 99: (0.1, 0.2, 0.9, 1.0)},
         "alpha": {20: 0.0, 50: 0.7, 80: 0.2}},
        {"name": "alpha_map_single", "pts": {0: (0.9, 0.2, 0.1, 1.0), 99: (0.1, 0.2, 0.9, 1.0)}, "alpha": {50: 0.7}},
        {"name": "alpha_map_out_of_range", "pts": {10: (0.3, 0.3, 0.3, 0.3), 60: (0.6, 0.6, 0.6, 0.6)},
         "alpha": {-10: 0.5, 40: 0.1, 99: 0.9}},
    ]
    for k in range(12):
        n = int(rng.integers(2, 9))
        pos = rng.choice(np.arange(-10, 110), size=n, replace=False)
        pts = {int(p): tuple(float(x) for x in np.round(rng.uniform(0, 1, 4), 4)) for p in pos}
        case = {"name": f"random_{k}", "pts": pts}
        if k % 3 == 0:
            apos = rng.choice(np.arange(0, 100), size=int(rng.integers(2, 5)), replace=False)
            case["alpha"] = {int(p): float(np.round(rng.uniform(0, 1), 4)) for p in apos}
        cases.append(case)
    # the control points the tests and the benchmark use
    import sys
    sys.path.insert(0, str(HERE.parent.parent))
    from monkeygl_amd import workloads as W
    for s in (0, 11, 12, 13, 14, 100, 101, 102, 103, 104):
        cases.append({"name": f"workload_seed_{s}", "pts": W.tf_points(s)})
    return cases


def render_info_cases():
    return [
        {"name": "rotate_z_steps", "ops": [("rotate", 30.0, 0.0)] * 12},
        {"name": "rotate_x_steps", "ops": [("rotate", 0.0, 45.0)] * 8},
        {"name": "mixed", "ops": [("rotate", 20.0, 35.0), ("zoom", 1.7), ("pan", 13.0, -21.0), ("rotate", -75.0, 110.0),
                                  ("zoom", 0.6), ("rotate", 5.5, -3.25)]},
        {"name": "presets", "ops": [("zoom", 2.0), ("preset", 0), ("preset", 1), ("zoom", 1.5), ("preset", 2),
                                    ("preset", 3), ("preset", 4), ("preset", 5), ("rotate", 10.0, 10.0)]},
        {"name": "many_small", "ops": [("rotate", 1.0, 0.5)] * 40},
    ]


def main():
    L = load()
    tf = []
    for c in tf_cases():
        lut = ref_tf(L, c["pts"], c.get("alpha"))
        tf.append({"name": c["name"], "pts": {str(k): list(v) for k, v in c["pts"].items()},
                   "alpha": {str(k): v for k, v in c.get("alpha", {}).items()},
                   "lut": None if lut is None else lut.tolist()})
    (HERE / "tf_luts.json").write_text(json.dumps(tf))
    ri = []
    for c in render_info_cases():
        ri.append({"name": c["name"], "ops": [list(o) for o in c["ops"]], "states": ref_render_info_run(L, c["ops"])})
    sizes = [(w, h, *ref_vr_size(L, w, h)) for (w, h) in [(512, 512), (768, 768), (1024, 1024), (1537, 1024),
                                                            (2048, 2048), (333, 101), (1600, 900), (3000, 700)]]
    (HERE / "render_info.json").write_text(json.dumps({"sequences": ri, "vr_sizes": sizes}))
    geo = cpr_geometry_cases(L)
    (HERE / "cpr_geometry.json").write_text(json.dumps(geo))
    mpr = mpr_geometry_cases(L)
    (HERE / "mpr_geometry.json").write_text(json.dumps(mpr))
    print(f"wrote {len(tf)} TF cases, {len(ri)} RenderInfo sequences, {len(sizes)} VR sizes, "
          f"{sum(len(v) for v in geo.values())} CPR and {sum(len(v) for v in mpr.values())} MPR geometry vectors")


if __name__ == "__main__":
    main()
