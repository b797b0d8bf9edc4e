"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, always as the checker / baseline and never as the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"


class OracleScene(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int), ("vertices", C.c_void_p), ("d_vertices", C.c_void_p),
        ("n_tris", C.c_int), ("faces", C.c_void_p), ("tri_uv", C.c_void_p), ("tri_obj", C.c_void_p),
        ("n_objects", C.c_int), ("obj_mat", C.c_void_p), ("obj_albedo", C.c_void_p), ("obj_rough", C.c_void_p),
        ("obj_light", C.c_void_p),
        ("n_texels", C.c_int), ("texels", C.c_void_p), ("d_texels", C.c_void_p),
        ("n_lights", C.c_int), ("light_emit", C.c_void_p), ("d_light_emit", C.c_void_p), ("light_obj", C.c_void_p),
        ("obj_normal", C.c_void_p),
    ]


class OracleCamera(C.Structure):
    _fields_ = [("look_from", C.c_float * 3), ("look_at", C.c_float * 3), ("up", C.c_float * 3),
                ("vfov", C.c_float), ("height", C.c_int), ("width", C.c_int), ("kind", C.c_int),
                ("view_width", C.c_float), ("d_params", C.c_float * 10)]


class OracleConfig(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "n_pass", "spp_interior", "enable_light_visable", "spp_primary_edge", "spp_secondary_edge", "max_bounce",
        "mis_light_samples", "mis_bsdf_samples", "primary_edge_preprocess_rounds",
        "secondary_edge_preprocess_rounds", "primary_edge_hypercube_resolution")] + [
        ("secondary_edge_hypercube_resolution", C.c_uint32 * 3)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
        _lib = C.CDLL(str(LIB_PATH))
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class Scene:
    """Host copy of a FlatScene (torch_psdr_amd.flat) plus optional tangents."""

    def __init__(self, flat, d_vertices=None, d_texels=None, d_emit=None):
        f = lambda t: np.ascontiguousarray(t.detach().cpu().numpy())
        self.vertices = f(flat.vertices).astype(np.float32)
        self.faces = f(flat.faces).astype(np.int32)
        self.tri_uv = f(flat.tri_uv).astype(np.float32)
        self.tri_obj = f(flat.tri_obj).astype(np.int32)
        self.texels = f(flat.texels).astype(np.float32)
        self.emit = f(flat.light_emit).astype(np.float32).reshape(-1)
        objs = flat.objects
        self.obj_mat = np.array([{0: 0, 1: 1, 2: 2, 3: 3, 4: 4}[o["material"]] for o in objs], dtype=np.int32)
        self.obj_albedo = np.array([[o["albedo_offset"], o["albedo_h"], o["albedo_w"]] for o in objs],
                                   dtype=np.int32).reshape(-1)
        self.obj_rough = np.array([[o["rough_offset"], o["rough_h"], o["rough_w"]] for o in objs],
                                  dtype=np.int32).reshape(-1)
        self.obj_light = np.array([o["light"] for o in objs], dtype=np.int32)
        self.obj_normal = np.array([[o.get("normal_offset", -1), o.get("normal_h", 0), o.get("normal_w", 0)]
                                    for o in objs], dtype=np.int32).reshape(-1)
        self.light_obj = np.array(flat.light_obj, dtype=np.int32)
        cv = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
        self.d_vertices, self.d_texels, self.d_emit = cv(d_vertices), cv(d_texels), cv(d_emit)
        s = OracleScene()
        s.n_vertices = self.vertices.shape[0]
        _a = lambda a: None if a is None else a.ctypes.data
        s.vertices, s.d_vertices = _a(self.vertices), _a(self.d_vertices)
        s.n_tris = self.faces.shape[0]
        s.faces, s.tri_uv, s.tri_obj = _a(self.faces), _a(self.tri_uv), _a(self.tri_obj)
        s.n_objects = len(objs)
        s.obj_mat, s.obj_albedo, s.obj_rough, s.obj_light = (_a(self.obj_mat), _a(self.obj_albedo),
                                                             _a(self.obj_rough), _a(self.obj_light))
        s.n_texels = self.texels.size
        s.texels, s.d_texels = _a(self.texels), _a(self.d_texels)
        s.n_lights = self.light_obj.size
        s.light_emit, s.d_light_emit, s.light_obj = _a(self.emit), _a(self.d_emit), _a(self.light_obj)
        s.obj_normal = _a(self.obj_normal)
        self.s = s


def camera(params, d_params=None):
    """params = look_from(3), look_at(3), up(3), vfov | view_width, height, width[, kind];
    d_params: tangent direction of the first 10 (renderD JVP w.r.t. the camera)."""
    c = OracleCamera()
    for i in range(3):
        c.look_from[i], c.look_at[i], c.up[i] = params[i], params[3 + i], params[6 + i]
    c.kind = int(params[12]) if len(params) > 12 else 0
    if c.kind == 1:
        c.view_width = params[9]
    else:
        c.vfov = params[9]
    c.height, c.width = int(params[10]), int(params[11])
    if d_params is not None:
        for i in range(10):
            c.d_params[i] = float(d_params[i])
    return c


def config(cfg):
    c = OracleConfig()
    names = [f[0] for f in OracleConfig._fields_][:11]
    for n, v in zip(names, cfg[:11]):
        setattr(c, n, int(v))
    for i in range(3):
        c.secondary_edge_hypercube_resolution[i] = int(cfg[11 + i])
    return c


def set_threads(n: int):
    load().oracle_set_threads(int(n))


def set_fd_measure(on: bool):
    """Boundary samples in the raster measure of a pixel mean (for finite-difference checks)."""
    load().oracle_set_fd_measure(int(bool(on)))


def diffuse_bsdf(color, sigma, n, v, l) -> np.ndarray:
    """DiffuseBsdf value (the oracle's restatement of src/diffuse_bsdf.cc:57-89)
    for arrays of cases: color (k,3), sigma (k,), shading normal, direction to
    the viewer and to the light (k,3 each) -> (k,3)."""
    color, n, v, l = (np.asarray(a, np.float32).reshape(-1, 3) for a in (color, n, v, l))
    sigma = np.asarray(sigma, np.float32).reshape(-1, 1)
    cases = np.ascontiguousarray(np.concatenate([color, sigma, n, v, l], 1), dtype=np.float32)
    out = np.zeros((cases.shape[0], 3), np.float32)
    load().oracle_diffuse_bsdf(_p(cases), cases.shape[0], _p(out))
    return out


def set_direct_unit_normal(on: bool):
    """Direct boundary term with the receiver's shadow-line direction and normal
    normalised and the shadow point's motion projected on that normal by a dot
    product (for finite-difference checks; off = the reference's formula)."""
    load().oracle_set_direct_unit_normal(int(bool(on)))


def pcg_kat(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    load().oracle_pcg_kat(C.c_uint64(initstate), C.c_uint64(initseq), n, _p(out))
    return out


def lane_draws(lane0: int, n_lanes: int, n_draws: int) -> np.ndarray:
    out = np.zeros((n_lanes, n_draws), dtype=np.uint32)
    load().oracle_lane_draws(C.c_uint32(lane0), n_lanes, n_draws, _p(out))
    return out


def lane_seed(lane0: int, n_lanes: int):
    st = np.zeros(n_lanes, dtype=np.uint64)
    inc = np.zeros(n_lanes, dtype=np.uint64)
    load().oracle_lane_seed(C.c_uint32(lane0), n_lanes, _p(st), _p(inc))
    return st, inc


def sincos(x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    load().oracle_sincos(_p(x), x.size, _p(s), _p(c))
    return s, c


def camera_matrices(params):
    cam = camera(params)
    outs = [np.zeros((4, 4), np.float32) for _ in range(4)]
    pos = np.zeros(3, np.float32)
    load().oracle_camera_matrices(C.byref(cam), *[_p(o) for o in outs], _p(pos))
    return outs, pos


def camera_rays(params, spp: int, d_params=None):
    """Camera rays of every lane; with d_params also the origins' tangents."""
    cam = camera(params, d_params)
    N = spp * int(params[10]) * int(params[11])
    o = np.zeros((N, 3), np.float32)
    d = np.zeros((N, 3), np.float32)
    pix = np.zeros(N, np.int32)
    do = np.zeros((N, 3), np.float32) if d_params is not None else None
    load().oracle_camera_rays(C.byref(cam), C.c_uint32(spp), _p(o), _p(d), _p(pix), _p(do))
    return (o, d, pix) if d_params is None else (o, d, pix, do)


def point_rays(params, pts, d_params, dpts):
    """Rays through world points (origin, direction) and the origins' tangents."""
    cam = camera(params, d_params)
    pts = np.ascontiguousarray(pts, np.float32)
    dpts = np.ascontiguousarray(dpts, np.float32)
    o = np.zeros_like(pts)
    d = np.zeros_like(pts)
    do = np.zeros_like(pts)
    load().oracle_point_rays(C.byref(cam), _p(pts), _p(dpts), pts.shape[0], _p(o), _p(d), _p(do))
    return o, d, do


def render_lanes(scene: Scene, cfg, cam_params, lanes):
    """Interior radiance (n, 3) of the listed lanes of a renderC."""
    lanes = np.ascontiguousarray(lanes, np.int64)
    out = np.zeros((lanes.size, 3), np.float32)
    c, k = config(cfg), camera(cam_params)
    load().oracle_render_lanes(C.byref(scene.s), C.byref(c), C.byref(k), _p(lanes), C.c_int64(lanes.size), _p(out))
    return out


def sh9_eval(coeff, dirs):
    """SH9Light radiance (n, 3) of unit directions dirs (n, 3) for coeff (9, 3)."""
    coeff = np.ascontiguousarray(coeff, np.float32).reshape(27)
    dirs = np.ascontiguousarray(dirs, np.float32)
    out = np.zeros_like(dirs)
    load().oracle_sh9_eval(_p(coeff), _p(dirs), dirs.shape[0], _p(out))
    return out


def camera_tangents(params, d_params):
    """Tangents of camera2world, world2camera, raster2camera along d_params."""
    cam = camera(params, d_params)
    outs = [np.zeros((4, 4), np.float32) for _ in range(3)]
    load().oracle_camera_tangents(C.byref(cam), *[_p(o) for o in outs])
    return outs


def project_points(params, pts):
    cam = camera(params)
    pts = np.ascontiguousarray(pts, np.float32)
    xy = np.zeros((pts.shape[0], 2), np.int32)
    valid = np.zeros(pts.shape[0], np.int32)
    load().oracle_project_points(C.byref(cam), _p(pts), pts.shape[0], _p(xy), _p(valid))
    return xy, valid.astype(bool)


def prepare(scene: Scene):
    nt = scene.faces.shape[0]
    tri = np.zeros((nt, 7), np.float32)
    cap = 3 * nt + 1
    topo = np.zeros((cap, 6), np.int32)
    geom = np.zeros((cap, 16), np.float32)
    ne = load().oracle_prepare(C.byref(scene.s), _p(tri), _p(topo), _p(geom), cap)
    return tri, topo[:ne], geom[:ne]


def samplers(scene: Scene, n_edges: int):
    no = scene.obj_mat.size
    nl = scene.light_obj.size
    lp, lc = np.zeros(no, np.float32), np.zeros(no, np.float32)
    ep, ec = np.zeros(max(n_edges, 1), np.float32), np.zeros(max(n_edges, 1), np.float32)
    pw, ar = np.zeros(max(nl, 1), np.float32), np.zeros(max(nl, 1), np.float32)
    load().oracle_samplers(C.byref(scene.s), _p(lp), _p(lc), _p(ep), _p(ec), _p(pw), _p(ar))
    return lp, lc, ep[:n_edges], ec[:n_edges], pw[:nl], ar[:nl]


def closest_hit(tri9: np.ndarray, ro: np.ndarray, rd: np.ndarray, brute: bool = False):
    tri9 = np.ascontiguousarray(tri9, np.float32).reshape(-1, 9)
    ro = np.ascontiguousarray(ro, np.float32)
    rd = np.ascontiguousarray(rd, np.float32)
    n = ro.shape[0]
    tri = np.zeros(n, np.int32)
    t, u, v = (np.zeros(n, np.float32) for _ in range(3))
    load().oracle_closest_hit(_p(tri9), C.c_int(tri9.shape[0]), _p(ro), _p(rd), C.c_int(n), C.c_int(int(brute)), _p(tri), _p(t), _p(u), _p(v))
    return tri, t, u, v


def hit_classes(tri9: np.ndarray, ro: np.ndarray, rd: np.ndarray) -> np.ndarray:
    """Per ray: bit 0 = the closest and second-closest triangle distances lie
    within 4 ulp (near tie), bit 1 = the hit's barycentrics lie within 1e-6 of
    an edge (grazing)."""
    tri9 = np.ascontiguousarray(tri9, np.float32).reshape(-1, 9)
    ro = np.ascontiguousarray(ro, np.float32)
    rd = np.ascontiguousarray(rd, np.float32)
    cls = np.zeros(ro.shape[0], np.int32)
    load().oracle_hit_classes(_p(tri9), tri9.shape[0], _p(ro), _p(rd), ro.shape[0], _p(cls))
    return cls


def lbvh(tri9: np.ndarray):
    tri9 = np.ascontiguousarray(tri9, np.float32).reshape(-1, 9)
    n = tri9.shape[0]
    morton = np.zeros(n, np.uint32)
    order = np.zeros(n, np.int32)
    child = np.zeros((n - 1, 2), np.int32)
    parent = np.zeros(2 * n - 1, np.int32)
    box = np.zeros((n - 1, 12), np.float32)
    rc = load().oracle_lbvh(_p(tri9), n, _p(morton), _p(order), _p(child), _p(parent), _p(box))
    assert rc == 0
    return morton, order, child, parent, box


def render(scene: Scene, cfg, cam_params, mode: int = 0, boundary: bool = True, lane_L: bool = False,
           d_camera=None):
    """mode 0: renderC image.  mode 1: (image, d_interior, d_primary, d_direct) JVP images
    along the scene's d_* tangents and the camera tangent d_camera (10 floats)."""
    H, W = int(cam_params[10]), int(cam_params[11])
    img = np.zeros((H, W, 3), np.float32)
    d_int = np.zeros_like(img) if mode == 1 else None
    d_pri = np.zeros_like(img) if (mode == 1 and boundary) else None
    d_dir = np.zeros_like(img) if (mode == 1 and boundary) else None
    N = int(cfg[1]) * H * W
    lL = np.zeros((N, 3), np.float32) if lane_L else None
    c, k = config(cfg), camera(cam_params, d_camera)
    load().oracle_render(C.byref(scene.s), C.byref(c), C.byref(k), mode, _p(img), _p(d_int), _p(d_pri), _p(d_dir),
                         _p(lL))
    if mode == 0:
        return (img, lL) if lane_L else img
    return img, d_int, d_pri, d_dir
