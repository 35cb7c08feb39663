"""Pins of the CPU oracle against closed forms (no GPU).

Each test fixes one part of the oracle independently of its own code, chosen
so that a plausible mistake (dropped term, wrong sign or index, swapped axis,
double instead of float lattice) changes the expected value.
"""
import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd._capi import (MK_MPR_AVERAGE, MK_MPR_MINIP, MK_MPR_MIP, MK_RENDER_MIP, MK_RENDER_SURFACE,
                                MK_RENDER_VR, MK_SAMPLE_LINEAR, MK_SAMPLE_NEAREST, PlaneParams)
from oracle import oracle as O

K = 32768.0 / 32767.0


def ramp(dims, a, b, c, d):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (a * x + b * y + c * z + d).astype(np.int16)


# ---- samplers ---------------------------------------------------------------

def test_tex3_linear_is_exact_on_linear_ramps():
    dims = (20, 16, 12)
    v = ramp(dims, 3, -5, 7, 11)
    rng = np.random.default_rng(0)
    for _ in range(200):
        # texel coordinates well inside (no clamping): f = (t + 0.5) / N
        t = rng.uniform(0.0, np.array(dims) - 1.0)
        f = (t + 0.5) / np.array(dims)
        exact = (3 * t[0] - 5 * t[1] + 7 * t[2] + 11) / 32767.0
        assert abs(O.tex3_linear(v, *f) - exact) < 1e-12


def test_tex3_linear_voxel_centres_and_clamp():
    dims = (6, 5, 4)
    v = np.random.default_rng(1).integers(-1000, 1000, size=dims[::-1]).astype(np.int16)
    for z in range(4):
        for y in range(5):
            for x in range(6):
                f = ((x + 0.5) / 6, (y + 0.5) / 5, (z + 0.5) / 4)
                assert abs(O.tex3_linear(v, *f) - v[z, y, x] / 32767.0) < 1e-12
    # far outside clamps to the corner voxel
    assert abs(O.tex3_linear(v, -3.0, -3.0, -3.0) - v[0, 0, 0] / 32767.0) < 1e-12
    assert abs(O.tex3_linear(v, 4.0, 4.0, 4.0) - v[-1, -1, -1] / 32767.0) < 1e-12
    # -32768 reads as -1 (normalized-float clamp)
    w = np.full((2, 2, 2), -32768, np.int16)
    assert O.tex3_linear(w, 0.5, 0.5, 0.5) == -1.0


def test_tex1_linear_lut():
    n = 100
    lut = np.zeros((n, 4), np.float32)
    lut[:, 0] = np.arange(n)
    lut[:, 3] = np.arange(n) / 100.0
    for s in [0.0, 0.004, 0.005, 0.123, 0.5, 0.995, 1.0]:
        t = min(max(s * n - 0.5, 0.0), n - 1.0)
        got = O.tex1_linear(lut, s)
        assert abs(got[0] - t) < 1e-9
    assert abs(O.tex1_linear(lut, -5.0)[0] - 0.0) < 1e-12
    assert abs(O.tex1_linear(lut, float("nan"))[0] - 0.0) < 1e-12


# ---- VR ---------------------------------------------------------------------

def _vr(vol, lut, labels=None, ang=(0.0, 0.0), size=32, rt=MK_RENDER_VR, dims=None, spacing=(1.0, 1.0, 1.0),
        **kw):
    dims = dims or vol.shape[::-1]
    info = P.RenderInfo()
    info.set_vr_size(size, size)
    info.rotate(*ang)
    p = P.render_params(dims, spacing, info, labels or {0: P.LabelState(1.0, 400.0, 40.0)}, render_type=rt,
                        **kw)
    return p, O.render(p, vol, {0: lut}, precision="f64")


def _flat_lut(rgb, a):
    lut = np.zeros((100, 4), np.float32)
    lut[:, :3] = rgb
    lut[:, 3] = a
    return lut


def _saturating_alpha(alpha, exit_a=0.995, cap=1.0):
    """Accumulated alpha when every sample is taken at the 1/8 step (alpha
    corrected to 1-(1-a)^(1/8)) until the exit threshold or the label cap."""
    a8 = 1.0 - (1.0 - alpha) ** 0.125
    A = 0.0
    while A <= exit_a and A < cap:
        A += (1.0 - A) * a8
    return A


@pytest.mark.parametrize("alpha", [0.3, 0.5, 0.9])
def test_vr_constant_volume_ambient_only(alpha):
    """Constant volume: zero gradient, so only the 0.35 ambient term; every
    sample refines to the 1/8 step with corrected alpha; the ray saturates."""
    vol = np.full((24, 24, 24), 100, np.int16)
    rgb = np.array([0.9, 0.5, 0.2], np.float32)
    p, img = _vr(vol, _flat_lut(rgb, alpha))
    A = _saturating_alpha(float(np.float32(alpha)))
    exp = np.floor(np.minimum(1.0, 0.35 * rgb.astype(np.float64) * A) * 255).astype(int)
    c = img[16, 16].astype(int)
    assert np.all(np.abs(c - exp) <= 1), (c, exp)


def test_vr_label_alpha_cap():
    vol = np.full((24, 24, 24), 100, np.int16)
    rgb = np.array([1.0, 1.0, 1.0], np.float32)
    labels = {0: P.LabelState(0.5, 400.0, 40.0)}
    p, img = _vr(vol, _flat_lut(rgb, 0.2), labels=labels)
    A = _saturating_alpha(float(np.float32(0.2)), cap=0.5)
    exp = int(np.floor(0.35 * A * 255))
    assert abs(int(img[16, 16, 0]) - exp) <= 1


@pytest.mark.parametrize("sign", [1, -1])
def test_vr_gradient_shading_sign(sign):
    """Ramp along +y (the view direction at identity): the normal is +-y and
    the headlight is +y, so shading is 0.35+0.8+0.16 for +1 and 0.35 for -1.
    The VOI keeps samples away from the faces, where the one-voxel taps clamp
    to the same edge voxel and the normal vanishes."""
    dims = (24, 40, 24)
    vol = ramp(dims, 0, sign * 5, 0, 0)
    rgb = np.array([0.3, 0.3, 0.3], np.float32)
    info = P.RenderInfo()
    info.set_vr_size(32, 32)
    p = P.render_params(dims, (1.0, 1.0, 1.0), info, {0: P.LabelState(1.0, 1e6, 0.0)})
    p.voi.posterior, p.voi.anterior = 10, 30
    img = O.render(p, vol, {0: _flat_lut(rgb, 0.6)})
    A = _saturating_alpha(float(np.float32(0.6)))
    k = 1.31 if sign > 0 else 0.35
    exp = int(np.floor(min(1.0, k * 0.3 * A) * 255))
    assert abs(int(img[16, 16, 0]) - exp) <= 1, (img[16, 16], exp)


@pytest.mark.parametrize("gx,gy", [(8, 6), (7, 24)])
def test_vr_gradient_shading_oblique(gx, gy):
    """Ramp v = gx*x + gy*y: the one-voxel central differences give the
    normal (gx, gy, 0)/|g| exactly, the headlight at identity is +y, so
    d = gy/|g| (0.6 and 0.96) and every sample is shaded with
    k = 0.35 + 0.8 d + 0.16 d^8 (core/kernel.cu:69-77). At d = 0.96 the
    specular exponent moves the pixel by >= 3 counts from d^2 or d^4, at
    d = 0.6 the diffuse coefficient by 12; the bar is 1."""
    dims = (40, 40, 24)
    vol = ramp(dims, gx, gy, 0, 0)
    rgb = np.array([0.7, 0.7, 0.7], np.float32)
    info = P.RenderInfo()
    info.set_vr_size(32, 32)
    p = P.render_params(dims, (1.0, 1.0, 1.0), info, {0: P.LabelState(1.0, 1e6, 0.0)})
    p.voi.left, p.voi.right, p.voi.posterior, p.voi.anterior = 10, 30, 10, 30
    img = O.render(p, vol, {0: _flat_lut(rgb, 0.6)})
    A = _saturating_alpha(float(np.float32(0.6)))
    d = gy / np.hypot(gx, gy)
    k = 0.35 + 0.8 * d + 0.16 * d ** 8
    def pixel(kk):
        return int(np.floor(min(1.0, kk * float(rgb[0]) * A) * 255))

    exp = pixel(k)
    # the pin tells the rule from its plausible mistakes
    if gy == 24:
        assert abs(pixel(0.35 + 0.8 * d + 0.16 * d ** 2) - exp) > 1
        assert abs(pixel(0.35 + 0.8 * d + 0.16 * d ** 4) - exp) > 1
    else:
        assert abs(pixel(0.35 + 0.8 + 0.16 * d ** 8) - exp) > 1
    assert abs(int(img[16, 16, 0]) - exp) <= 1, (img[16, 16], exp)


def test_vr_empty_tf_background():
    vol = np.full((16, 16, 16), 500, np.int16)
    p, img = _vr(vol, np.zeros((100, 4), np.float32), background=(0.2, 0.4, 0.6, 1.0))
    assert np.all(img[..., 0] == int(0.2 * 255)) and np.all(img[..., 1] == int(np.float32(0.4) * 255))


def test_vr_float_lattice():
    """The sample lattice is float accumulation of the base step
    (kernel.cu:169, 213, 238), not k * step: on a deep volume the two put the
    last in-VOI sample of a MIP ray at visibly different ramp values."""
    nx, ny, nz = 8, 16, 2000
    dims = (nx, ny, nz)
    vol = ramp(dims, 0, 1000, 0, -8000)
    info = P.RenderInfo()
    info.set_vr_size(8, 8)
    p = P.render_params(dims, (1.0, 1.0, 1.0), info, {0: P.LabelState(1.0, 2550.0, 0.0)}, render_type=MK_RENDER_MIP)
    p.voi.anterior = ny // 2  # last in-VOI y index is in the interior
    img = O.render(p, vol, {0: _flat_lut([1, 1, 1], 1)}, precision="f64")
    # the same ray traced in numpy on both lattices
    u = np.float32((4 - 4.0) / 8.0)
    mls_y = np.float32(np.float32(nz) / np.float32(ny))
    L1 = np.float32(np.float32(1.0) / np.float32(nz))

    def last_value(acc_seq):
        best = -np.inf
        for acc in acc_seq:
            fy = float(np.float32((float(acc) - 0.866) * 1.0))
            py = fy * float(mls_y) + 0.5
            iy = int(np.float32(py) * ny)
            px = float(u) * float(nz / nx) + 0.5
            if 0 <= iy <= ny // 2 and 0 <= int(px * nx) < nx:
                t = py * ny - 0.5
                t = min(max(t, 0.0), ny - 1.0)
                hu = (1000 * t - 8000) * K
                best = max(best, (hu - 0.0) / 2550.0 + 0.5)
        return best

    acc_f, seq_f = np.float32(0.0), []
    while float(acc_f) < 1.732:
        seq_f.append(acc_f)
        acc_f = np.float32(acc_f + L1)
    seq_d = [k * float(L1) for k in range(len(seq_f))]
    exp_f = int(np.floor(min(1, last_value(seq_f)) * 255))
    exp_d = int(np.floor(min(1, last_value(seq_d)) * 255))
    assert exp_f != exp_d  # the test discriminates the two lattices
    assert abs(int(img[4, 4, 0]) - exp_f) <= 1 and abs(int(img[4, 4, 0]) - exp_d) > 1


def test_mip_plateau_closed_form():
    vol = np.full((24, 24, 24), -1000, np.int16)
    vol[:, 8:16, :] = 600
    p, img = _vr(vol, _flat_lut([1, 1, 1], 1), rt=MK_RENDER_MIP, labels={0: P.LabelState(1.0, 2000.0, 0.0)})
    exp = int(np.floor(min(1.0, (600 * K) / 2000.0 + 0.5) * 255))
    assert int(img[16, 16, 0]) == exp


@pytest.mark.parametrize("alpha0,alpha1,expect", [(1.0, 0.0, "slab0"), (1.0, 0.3, "slab1"), (0.0, 0.3, "slab1"),
                                                   (0.0, 0.0, "background")])
def test_mip_ignores_labels_with_zero_alpha(alpha0, alpha1, expect):
    """MIP takes the windowed value only of samples whose label has alpha > 0
    (core/kernel.cu:334-343): a 600 HU slab of label 0 and a 1500 HU slab of
    label 1 on the ray; the maximum is the slab of the labels that count, and
    the background when none does."""
    dims = (24, 24, 24)
    vol = np.full(dims[::-1], -1000, np.int16)
    vol[:, 4:8, :] = 600
    vol[:, 14:18, :] = 1500
    mask = np.zeros(dims[::-1], np.uint8)
    mask[:, 14:18, :] = 1
    info = P.RenderInfo()
    info.set_vr_size(32, 32)
    labels = {0: P.LabelState(alpha0, 2000.0, 0.0), 1: P.LabelState(alpha1, 2000.0, 0.0)}
    p = P.render_params(dims, (1.0, 1.0, 1.0), info, labels, render_type=MK_RENDER_MIP, background=(0.2, 0.2, 0.2, 1.0))
    img = O.render(p, vol, {0: _flat_lut([1, 1, 1], 1), 1: _flat_lut([1, 1, 1], 1)}, mask_zyx=mask)
    want = {"slab0": int(np.floor(min(1.0, 600 * K / 2000.0 + 0.5) * 255)), "slab1": 255,
            "background": int(np.float32(0.2) * 255)}[expect]
    assert want in (204, 255, 51)
    assert int(img[16, 16, 0]) == want, (img[16, 16], want)


def test_surface_oblique_normal():
    """Ramp v = 8x + 6y + 300 (every in-VOI sample above 400 HU): the first
    sample's normal is (0.8, 0.6, 0), the headlight +y, so d = 0.6 and the
    gray is d + 0.05 = 0.65 -> 165 (core/kernel.cu:441-462); without the 0.05
    it would be 153, with the normal's sign flipped the ray finds no hit."""
    dims = (40, 40, 24)
    info = P.RenderInfo()
    info.set_vr_size(32, 32)
    for gy, want in [(6, 165), (-6, None)]:
        vol = ramp(dims, 8, gy, 0, 300 if gy > 0 else 700)
        p = P.render_params(dims, (1.0, 1.0, 1.0), info, {0: P.LabelState(1.0, 400.0, 40.0)},
                            render_type=MK_RENDER_SURFACE, background=(0.0, 0.0, 0.0, 1.0))
        p.voi.left, p.voi.right, p.voi.posterior, p.voi.anterior = 10, 30, 10, 30
        img = O.render(p, vol, {0: _flat_lut([1, 1, 1], 1)})
        if want is None:
            assert int(img[16, 16, 0]) == 0
        else:
            assert int(img[16, 16, 0]) == want and int(img[16, 16, 1]) == want, img[16, 16]


def test_surface_first_hit():
    """Half-space brighter than 400 HU entered along +y: the first hit's
    normal is +y, the headlight +y, so gray = 1 + 0.05 saturates to 255; rays
    that never reach 400 HU show the background."""
    dims = (24, 24, 24)
    vol = np.full(dims[::-1], 0, np.int16)
    vol[:, 12:, :] = 1000
    p, img = _vr(vol, _flat_lut([1, 1, 1], 1), rt=MK_RENDER_SURFACE, background=(0.0, 0.0, 0.0, 1.0))
    assert img[16, 16, 0] == 255
    vol2 = np.full(dims[::-1], 0, np.int16)
    p, img2 = _vr(vol2, _flat_lut([1, 1, 1], 1), rt=MK_RENDER_SURFACE, background=(0.0, 0.0, 0.0, 1.0))
    assert img2.max() == 0


# ---- MPR / CPR ----------------------------------------------------------------

def test_nearest_axial_is_a_raw_slice():
    dims = (20, 18, 10)
    vol = np.random.default_rng(3).integers(-1024, 3000, size=dims[::-1]).astype(np.int16)
    p = P.plane_params(dims, (1.0, 1.0, 1.0), P.PLANE_AXIAL, sampling=MK_SAMPLE_NEAREST)
    out = O.render_plane(p, vol, precision="f32")
    # fz = 1 - 0.5 (not inverted): slice floor(0.5 * nz)
    assert np.array_equal(out[:, :20], vol[5])


@pytest.mark.parametrize("mode", [MK_MPR_AVERAGE, MK_MPR_MIP, MK_MPR_MINIP])
def test_slab_of_z_ramp(mode):
    """z-ramp, axial slab of 5 slices: Average is the centre value (linear),
    MIP the top and MinIP the bottom slice of the slab."""
    dims = (16, 16, 40)
    vol = ramp(dims, 0, 0, 50, 0)
    p = P.plane_params(dims, (1.0, 1.0, 1.0), P.PLANE_AXIAL, mpr_type=mode, sampling=MK_SAMPLE_NEAREST,
                       thickness=5.0)
    out = O.render_plane(p, vol)
    zc = 20  # floor(0.5 * 40)
    # dir_n = dirH x dirV = +z in patient space, z is flipped: t>0 moves to lower z index
    exp = {MK_MPR_AVERAGE: 50 * zc, MK_MPR_MIP: 50 * (zc + 2), MK_MPR_MINIP: 50 * (zc - 2)}[mode]
    assert np.all(out[2:-2, 2:-2] == exp), (mode, out[8, 8], exp)


def test_mpr_out_of_volume_sentinel():
    dims = (10, 10, 10)
    vol = np.full(dims[::-1], 7, np.int16)
    p = P.plane_params(dims, (1.0, 1.0, 1.0), P.PLANE_AXIAL, mpr_type=MK_MPR_MIP, size=(30, 30))
    out = O.render_plane(p, vol)
    assert out[0, 0] == -32768 and out[15, 15] == 7


def test_cpr_straight_line_is_a_ramp_row():
    """Row y samples P_y + x*D_y in voxel units over dims: on a linear ramp
    that is the ramp at texel coordinates (f*N - 0.5), exactly."""
    dims = (32, 24, 16)
    vol = ramp(dims, 4, 3, 2, 100)
    h, w = 10, 20
    pts = np.array([[2.0, 4.0 + j, 7.5] for j in range(h)])
    dirs = np.tile([1.0, 0.0, 0.0], (h, 1))
    out = O.render_cpr(w, pts, dirs, invert_z=True, vol_zyx=vol)
    for j in range(h):
        for x in range(w):
            tx, ty, tz = 2.0 + x - 0.5, 4.0 + j - 0.5, 7.5 - 0.5
            inside = 0 <= 2.0 + x <= 32
            if inside and tx <= 31:
                exp = int((4 * max(tx, 0) + 3 * ty + 2 * tz + 100) * K)
                assert out[j, x] == exp, (j, x, out[j, x], exp)
    far = O.render_cpr(4, np.array([[100.0, 0.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]), True, vol)
    assert np.all(far == -2048)
