"""The demos' colour buffers and particle renderer without OpenGL, through the
C ABI: colours equal the reference's fill rules (fluidSystem.cpp:160-181,
poiseuilleFlowSystem.cpp:148-169, PeristalsisSystem.cpp:409-442) exactly; the
renderer's frame equals a CPU restatement of the demo's pipeline (the
modelview of DamBreak.Demo/main.cpp:109-114, gluPerspective, point sprites of
pointRadius * pointScale / distance pixels, the sphere shader of shaders.cpp,
GL_LESS depth in draw order), written here in FP32 step by step."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cmag_amd import _lib as L

pytestmark = pytest.mark.gpu

F = np.float32
RAMP = np.array([[1, 0, 0], [1, 0.5, 0], [1, 1, 0], [0, 1, 0], [0, 1, 1], [0, 0, 1], [1, 0, 1]], F)


def _hip():
    return C.CDLL("libamdhip64.so")


def _device_to_host(ptr, nbytes):
    out = np.empty(nbytes, np.uint8)
    rc = _hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(nbytes), 2)
    assert rc == 0
    return out


def _color_buffer(h, n):
    p = C.c_void_p()
    L.call("cmag_system_color_buffer", h, C.byref(p))
    return _device_to_host(p.value, 16 * n).view(np.float32).reshape(n, 4)


def color_ramp(t):
    t = (F(t) * F(6.0)).astype(F)
    i = np.floor(t).astype(int)
    u = (t - np.floor(t)).astype(F)
    rgb = (RAMP[i] + (u[:, None] * (RAMP[i + 1] - RAMP[i])).astype(F)).astype(F)
    return np.column_stack([rgb, np.ones(len(t), F)])


def _system(kind):
    s = C.c_void_p()
    if kind == "dambreak":
        L.call("cmag_dambreak_create", 32, 32, 1, 1, 128, 64, 4, 1.0 / 64, C.byref(s))
    elif kind == "poiseuille":
        L.call("cmag_poiseuille_create", 32, 64, 1, 3, 32, 128, 4, 1.0 / 128, C.byref(s))
    else:
        L.call("cmag_peristalsis_create", 1e-4, 32, 32, 1, 0.2, 1e-3, 0.1, 0.0, 0.0, 0.0, 3, 64, 64, 4,
               1.0 / 128, C.byref(s))
    L.call("cmag_system_reset", s)
    return s


def _params(s):
    p = L.CmagParams()
    L.call("cmag_system_get_params", s, C.byref(p))
    return p


def _positions(s, n):
    out = np.empty((n, 4), np.float32)
    L.call("cmag_system_get_array", s, L.ARRAY_POSITION, out.ctypes.data, 0, n)
    return out


@pytest.mark.parametrize("kind", ["dambreak", "poiseuille"])
def test_class_colours_follow_the_reference_fill(device, kind):
    s = _system(kind)
    n = L.lib().cmag_system_num_particles(s)
    p = _params(s)
    got = _color_buffer(s, n)
    i = np.arange(n)
    fluid = p.fluid_size[0] * p.fluid_size[1] * p.fluid_size[2]
    t = np.full(n, F(0.7), F)
    if kind == "dambreak":
        type_one = fluid + (p.grid_size[0] - p.boundary_offset + 1) + 2 * (p.grid_size[1] - p.boundary_offset)
        t[i < type_one] = F(0.2)
        t[i < fluid] = F(0.5)
    else:
        t[i < fluid] = F(0.5)
        t[(i % p.grid_size[0] == 0) & (i < fluid)] = F(0.2)
    assert np.array_equal(got, color_ramp(t))
    assert np.array_equal(got[0 if kind == "dambreak" else 1], [0, 1, 0, 1])  # colorRamp(0.5): green fluid
    # Coloring() is Peristalsis's only
    assert L.lib().cmag_system_coloring(s) == L.CMAG_E_INVALID
    L.call("cmag_system_destroy", s)


def test_peristalsis_coloring_marks_the_inlet_slab(device):
    s = _system("peristalsis")
    n = L.lib().cmag_system_num_particles(s)
    p = _params(s)

    def want(pos):
        c = np.empty((n, 4), F)
        c[:] = [255, 255, 255, 1]
        c[pos[:, 0] < F(p.world_origin[0]) + F(4.0) * F(p.particle_radius)] = [255, 0, 0, 1]
        c[pos[:, 3] != 0] = [0, 0, 0, 1]
        return c

    got = _color_buffer(s, n)
    w = want(_positions(s, n))
    assert np.array_equal(got, w)
    assert (w[:, 1] == 0).sum() > 0 and (w[:, 0] == 0).sum() > 0
    # positions move; Coloring() recolours from the new ones
    pos = _positions(s, n)
    fl = pos[:, 3] == 0
    pos[fl, 0] += F(3.0) * F(p.particle_radius)
    L.call("cmag_system_set_array", s, L.ARRAY_POSITION, pos.ctypes.data, 0, n)
    L.call("cmag_system_coloring", s)
    assert np.array_equal(_color_buffer(s, n), want(_positions(s, n)))
    L.call("cmag_system_destroy", s)


# ---- renderer --------------------------------------------------------------------
class View(C.Structure):
    _fields_ = [("translate", C.c_float * 3), ("rotate", C.c_float * 2), ("fov_deg", C.c_float),
                ("znear", C.c_float), ("zfar", C.c_float), ("point_radius", C.c_float),
                ("max_point_size", C.c_float), ("clear", C.c_float * 4)]


def render_cpu(view, pos, color, w, h):
    """the demo's pipeline in FP32, one operation at a time"""
    a, b = np.deg2rad(view.rotate[0]), np.deg2rad(view.rotate[1])
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    R = np.array([[cb, 0, sb], [sa * sb, ca, -sa * cb], [-ca * sb, sa, ca * cb]])
    m = np.zeros((3, 4), F)
    m[:, :3] = R.astype(F)
    m[:, 3] = np.array(view.translate[:], F)
    f = 1.0 / np.tan(0.5 * np.deg2rad(view.fov_deg))
    fx, fy = F(f / (w / h)), F(f)
    za = F((view.zfar + view.znear) / (view.znear - view.zfar))
    zb = F(2.0 * view.zfar * view.znear / (view.znear - view.zfar))
    scale = F(h / np.tan(0.5 * np.deg2rad(view.fov_deg)))
    radius, max_size = F(view.point_radius), F(view.max_point_size)
    p = pos.astype(F)
    e = [((m[r, 0] * p[:, 0] + m[r, 1] * p[:, 1]).astype(F) + m[r, 2] * p[:, 2]).astype(F) + m[r, 3]
         for r in range(3)]
    xe, ye, ze = (x.astype(F) for x in e)
    cw = -ze
    cx, cy, cz = fx * xe, fy * ye, (za * ze + zb).astype(F)
    vis = (cw > 0) & (np.abs(cx) <= cw) & (np.abs(cy) <= cw) & (np.abs(cz) <= cw)
    inv = (F(1.0) / cw).astype(F)
    xw = (((cx * inv) + F(1)) * F(0.5) * F(w)).astype(F)
    yw = (((cy * inv) + F(1)) * F(0.5) * F(h)).astype(F)
    zw = (((cz * inv) + F(1)) * F(0.5)).astype(F)
    dist = np.sqrt((xe * xe + ye * ye).astype(F) + ze * ze).astype(F)
    size = np.minimum(np.maximum(radius * (scale / dist), F(1)), max_size).astype(F)
    vis &= zw < 1
    best = np.full((h, w), np.inf)
    win = np.full((h, w), -1)
    for i in np.nonzero(vis)[0]:
        half = F(0.5) * size[i]
        x0, x1 = max(0, int(np.ceil(xw[i] - half - F(0.5)))), min(w - 1, int(np.ceil(xw[i] + half - F(0.5))) - 1)
        y0, y1 = max(0, int(np.ceil(yw[i] - half - F(0.5)))), min(h - 1, int(np.ceil(yw[i] + half - F(0.5))) - 1)
        if x1 < x0 or y1 < y0:
            continue
        px, py = np.meshgrid(np.arange(x0, x1 + 1), np.arange(y0, y1 + 1))
        sinv = F(1) / size[i]
        nx = F(2) * ((px.astype(F) + F(0.5) - xw[i]) * sinv)
        ny = F(2) * ((py.astype(F) + F(0.5) - yw[i]) * sinv)
        inside = (nx * nx + ny * ny) <= F(1)
        sel = inside & (zw[i] < best[py, px])  # GL_LESS in draw order
        best[py[sel], px[sel]] = zw[i]
        win[py[sel], px[sel]] = i
    img = np.empty((h, w, 4), np.uint8)
    img[:] = np.round(np.array(view.clear[:]) * 255).astype(np.uint8)
    for py, px in zip(*np.nonzero(win >= 0)):
        i = win[py, px]
        sinv = F(1) / size[i]
        nx = F(2) * ((F(px) + F(0.5) - xw[i]) * sinv)
        ny = F(2) * ((F(py) + F(0.5) - yw[i]) * sinv)
        nz = np.sqrt(max(F(1) - (nx * nx + ny * ny), F(0))).astype(F)
        d = max(F(0), (F(0.577) * nx + F(0.577) * ny) + F(0.577) * nz)
        c = np.clip(color[i] * F(d), 0, 1)
        img[h - 1 - py, px] = (c * F(255) + F(0.5)).astype(np.uint8)
    return img, win[::-1]


def test_renderer_matches_the_cpu_pipeline(device):
    import torch
    from tests.gpu_util import ptr, to_dev
    rng = np.random.default_rng(7)
    n = 600
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.6, 0.6, (n, 3))
    pos[:40, :3] = pos[40:80, :3] + rng.normal(0, 0.01, (40, 3))  # overlapping sprites
    pos[80] = [5, 0, 0, 0]  # outside the view volume
    color = rng.uniform(0, 1, (n, 4)).astype(np.float32)
    v = View()
    L.lib().cmag_render_view_default(C.byref(v))
    v.translate[0], v.translate[1], v.translate[2] = 0.1, -0.05, -2.2
    v.rotate[0], v.rotate[1] = 20.0, -35.0
    v.point_radius = 0.04
    v.clear[0], v.clear[1], v.clear[2], v.clear[3] = 1.0, 1.0, 1.0, 1.0
    w, h = 160, 120
    r = C.c_void_p()
    L.call("cmag_renderer_create", w, h, C.byref(r))
    dpos, dcol = to_dev(pos), to_dev(color)
    L.call("cmag_renderer_display", r, C.byref(v), ptr(dpos), ptr(dcol), n, None)
    torch.cuda.synchronize()
    got = np.empty((h, w, 4), np.uint8)
    L.call("cmag_renderer_read", r, got.ctypes.data, None)
    want, win = render_cpu(v, pos[:, :3], color, w, h)
    covered = win >= 0
    assert covered.mean() > 0.1
    same = np.all(got == want, axis=2)
    close = np.all(np.abs(got.astype(int) - want.astype(int)) <= 1, axis=2)
    print(f"covered {covered.mean():.3f}, identical pixels {same.mean():.4f}, within 1: {close.mean():.4f}")
    # a pixel could differ only where two sprites' depths or a disc edge sit
    # within rounding of each other (none does in this scene)
    assert same.mean() >= 0.999 and close.mean() >= 0.999
    # no colour: white spheres, brightest towards the light (up, right, front)
    L.call("cmag_renderer_display", r, C.byref(v), ptr(dpos), None, n, None)
    white = np.empty_like(got)
    L.call("cmag_renderer_read", r, white.ctypes.data, None)
    assert np.all(white[..., 0] == white[..., 1]) and np.all(white[..., 1] == white[..., 2])
    L.call("cmag_renderer_destroy", r)
    del dpos, dcol


def test_single_sprite_size_and_shading(device):
    """one particle straight ahead: a disc of radius * (h / tan(30 deg)) / d pixels
    across, centred, lit brightest on its upper right"""
    from tests.gpu_util import ptr, to_dev
    w = h = 101
    v = View()
    L.lib().cmag_render_view_default(C.byref(v))
    v.translate[0], v.translate[1], v.translate[2] = 0.0, 0.0, -2.0
    v.point_radius = 0.2
    v.clear[0], v.clear[1], v.clear[2], v.clear[3] = 0.0, 0.0, 1.0, 1.0
    pos = np.array([[0, 0, 0, 0]], np.float32)
    r = C.c_void_p()
    L.call("cmag_renderer_create", w, h, C.byref(r))
    d = to_dev(pos)
    L.call("cmag_renderer_display", r, C.byref(v), ptr(d), None, 1, None)
    img = np.empty((h, w, 4), np.uint8)
    L.call("cmag_renderer_read", r, img.ctypes.data, None)
    size = 0.2 * (h / np.tan(np.deg2rad(30.0))) / 2.0
    lit = ~np.all(img == [0, 0, 255, 255], axis=2)  # the sprite, its unlit side included
    area = lit.sum()
    assert abs(area - np.pi * (size / 2) ** 2) < 0.15 * np.pi * (size / 2) ** 2, (area, size)
    ys, xs = np.nonzero(lit)
    assert abs(xs.mean() - 50) < 1 and abs(ys.mean() - 50) < 1
    k = np.argmax(img[..., 0].astype(int) + lit)
    ky, kx = divmod(k, w)
    assert kx > 50 and ky < 50  # top row first: the light comes from the upper right
    L.call("cmag_renderer_destroy", r)
    del d


RENDER = Path(__file__).resolve().parent.parent / "cmag_amd" / "examples" / "dambreak_render"


def test_render_example_writes_the_demo_frames(device, tmp_path):
    assert RENDER.exists(), "build the examples first (python -m cmag_amd.build)"
    proc = subprocess.run([str(RENDER), str(tmp_path), "2", "20", "320", "256"], capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 0, proc.stderr
    for k in range(2):
        raw = (tmp_path / f"frame{k}.ppm").read_bytes()
        head = b"P6\n320 256\n255\n"
        assert raw.startswith(head)
        img = np.frombuffer(raw[len(head):], np.uint8).reshape(256, 320, 3)
        green = (img[..., 1] > 60) & (img[..., 0] < 20) & (img[..., 2] < 20)  # colorRamp(0.5), lit
        assert green.sum() > 500
        assert (img == 255).all(axis=2).mean() > 0.3  # white background
