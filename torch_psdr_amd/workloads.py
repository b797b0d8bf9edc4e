"""Synthetic scenes of the benchmark configurations, built through the drop-in API.

  C0  icosphere level 3 (r=1, origin), 10x10 ground quad at y=-1, 2x2 quad light
      at y=4 with radiance 10, albedo 0.5, camera (0,1,5) -> origin, vfov 45 deg,
      256x256, 4 spp, max_bounce 2
  C1  C0 + a 256x256 UV-sphere-mapped texture on the sphere, texels U[0.2, 0.8]
      (seed 1); target texture seed 2; 8 spp, depth 2
  C2  icosphere level 5 (20480 tris), vertices radially perturbed by N(0, 0.02)
      (seed 3), 512x512 texture, ground + light as C0; 512x512, 16 spp, depth 3,
      spp_primary_edge = spp_secondary_edge = 4 (1M boundary samples)
  C3  64 scenes, each a geodesic sphere of N = 20 k^2 triangles with log10 N
      uniform on [3, 5.3] (seed 4), randomly scaled in [0.5, 2] and rotated,
      vertices radially perturbed by N(0, 0.02), a 256x256 texture, and the C0
      ground + light; resolution uniform in {128, 256, 512, 1024}^2, spp in
      {4, 16, 64}, depth 3; target = the unperturbed sphere with a second
      texture; all scenes share one set of per-render buffers
  C4  1024 x 1024 height-grid terrain (2 * 1023^2 tris), heights from 6-octave
      fBm value noise (amplitude 1, seed 5), 1024^2 texture (seed 1), two quad
      area lights; 1024x1024, 64 spp, depth 4; gradients w.r.t. the heights and
      texels
Geometry is generated with numpy (float64, then rounded to fp32 once).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch


def icosphere(level: int, radius: float = 1.0):
    """Subdivided icosahedron: (V, 3) float32 vertices, (F, 3) int32 faces."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = (np.stack(verts) * radius).astype(np.float32)
    return V, np.array(faces, dtype=np.int32)


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    F = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    return V / np.linalg.norm(V, axis=1, keepdims=True), F


def geodesic_sphere(k: int, radius: float = 1.0):
    """Class-I geodesic sphere of frequency k: every icosahedron face split into
    k^2 triangles (20 k^2 in all, wound as the face), grid points projected to
    the sphere.  A grid point shared by faces (on an icosahedron edge or vertex)
    is one vertex.  (V, 3) float32, (F, 3) int32."""
    base, faces = _icosahedron()
    I, J = np.meshgrid(np.arange(k + 1), np.arange(k + 1), indexing="ij")
    m = I + J <= k
    i, j = I[m], J[m]
    npts = i.size
    local = np.full((k + 2, k + 2), -1, dtype=np.int64)
    local[i, j] = np.arange(npts)
    # a grid point's key: its nonzero (corner vertex, weight) pairs, sorted
    big = 12 * (k + 1)
    w = np.stack([k - i - j, i, j], 1)                       # (npts, 3)
    keys = []
    pos = []
    for a, b, c in faces:
        ids = np.array([a, b, c])
        code = np.where(w > 0, ids[None, :] * (k + 1) + w, big)
        keys.append(np.sort(code, axis=1))
        pos.append((base[a][None, :] * w[:, :1] + base[b][None, :] * w[:, 1:2] + base[c][None, :] * w[:, 2:]) / k)
    keys = np.concatenate(keys)
    pos = np.concatenate(pos)
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    V = pos[first]
    V = V / np.linalg.norm(V, axis=1, keepdims=True) * radius
    ii, jj = I[I + J < k], J[I + J < k]
    up = np.stack([local[ii, jj], local[ii + 1, jj], local[ii, jj + 1]], 1)
    di, dj = I[I + J < k - 1], J[I + J < k - 1]
    down = np.stack([local[di + 1, dj], local[di + 1, dj + 1], local[di, dj + 1]], 1)
    tri = np.concatenate([up, down])
    F = np.concatenate([inv[tri + f * npts] for f in range(faces.shape[0])])
    return V.astype(np.float32), F.astype(np.int32)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """A uniformly random 3x3 rotation (QR of a Gaussian matrix, signs fixed)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def sphere_uv(V: np.ndarray) -> np.ndarray:
    """UV-sphere map of unit-sphere directions: u = 0.5 + atan2(z, x)/2pi, v = 0.5 + asin(y)/pi."""
    v = V.astype(np.float64)
    v = v / np.linalg.norm(v, axis=1, keepdims=True)
    u = 0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * math.pi)
    w = 0.5 + np.arcsin(np.clip(v[:, 1], -1, 1)) / math.pi
    return np.stack([u, w], 1).astype(np.float32)


def quad(size: float, y: float, facing_down: bool = False):
    h = size / 2.0
    V = np.array([[-h, y, -h], [h, y, -h], [h, y, h], [-h, y, h]], dtype=np.float32)
    F = np.array([[0, 2, 1], [0, 3, 2]] if not facing_down else [[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    return V, F


@dataclass
class Workload:
    name: str
    scene: object
    integrator: object
    params: dict


def _t(x, device, requires_grad=False):
    return torch.tensor(x, device=device, requires_grad=requires_grad)


def camera(dr, device, height, width):
    return dr.PerspectiveCamera(
        look_from=torch.tensor([0.0, 1.0, 5.0], device=device),
        look_at=torch.tensor([0.0, 0.0, 0.0], device=device),
        up=torch.tensor([0.0, 1.0, 0.0], device=device),
        vfov=torch.tensor([math.pi / 4.0], device=device),
        height=height, width=width)


def texture(seed: int, res: int, device) -> torch.Tensor:
    g = torch.Generator().manual_seed(seed)
    return (0.2 + 0.6 * torch.rand((res, res, 3), generator=g, dtype=torch.float32)).to(device)


def build_scene(dr, device, *, level=3, res=256, tex=None, sphere_vertices=None, spp=4, max_bounce=2,
                spp_edge=1, vertices_grad=False, light_emit=10.0, albedo=0.5):
    """Sphere + ground quad + quad light, as the configurations describe."""
    Vs, Fs = icosphere(level)
    if sphere_vertices is not None:
        Vs = sphere_vertices
    vs = torch.tensor(Vs, device=device, requires_grad=vertices_grad)
    fs = torch.tensor(Fs, device=device)
    if tex is None:
        mat = dr.DiffuseMaterial(torch.tensor([albedo] * 3, device=device))
        sphere = dr.Mesh(vs, None, fs, None, mat)
    else:
        uv = torch.tensor(sphere_uv(icosphere(level)[0]), device=device)
        sphere = dr.Mesh(vs, uv, fs, fs, dr.DiffuseMaterial(tex))
    Vg, Fg = quad(10.0, -1.0)
    ground = dr.Mesh(torch.tensor(Vg, device=device), None, torch.tensor(Fg, device=device), None,
                     dr.DiffuseMaterial(torch.tensor([albedo] * 3, device=device)))
    Vl, Fl = quad(2.0, 4.0, facing_down=True)
    light = dr.AreaLight(torch.tensor(Vl, device=device), torch.tensor(Fl, device=device),
                         torch.tensor([light_emit] * 3, device=device))
    cam = camera(dr, device, res, res)
    scene = dr.Scene([cam], [sphere, ground], [light])
    integ = dr.PathIntegrator(spp_interior=spp, max_bounce=max_bounce, spp_primary_edge=spp_edge,
                              spp_secondary_edge=spp_edge)
    return scene, integ, (vs, tex)


def c0(dr, device, res=256, spp=4):
    scene, integ, _ = build_scene(dr, device, level=3, res=res, spp=spp, max_bounce=2)
    return Workload("c0_forward_256_4spp_d2", scene, integ, {})


def perturbed_sphere(level: int, seed: int, sigma: float = 0.02) -> np.ndarray:
    V, _ = icosphere(level)
    rng = np.random.default_rng(seed)
    r = 1.0 + sigma * rng.standard_normal(V.shape[0])
    return (V.astype(np.float64) * r[:, None]).astype(np.float32)


def c2_parts(dr, device, res=512, spp=16, spp_edge=4, level=5, tex_res=512):
    """C2: perturbed level-5 icosphere (seed 3) with a 512^2 texture (seed 1),
    target = unperturbed mesh with texture seed 2; gradients w.r.t. the sphere's
    vertices and texels."""
    V0, F = icosphere(level)
    Vp = perturbed_sphere(level, seed=3)
    uv = torch.tensor(sphere_uv(V0), device=device)
    faces = torch.tensor(F, device=device)
    Vg, Fg = quad(10.0, -1.0)
    Vl, Fl = quad(2.0, 4.0, facing_down=True)
    ground_v = torch.tensor(Vg, device=device)
    ground_f = torch.tensor(Fg, device=device)
    ground_c = torch.tensor([0.5] * 3, device=device)
    light_v = torch.tensor(Vl, device=device)
    light_f = torch.tensor(Fl, device=device)
    emit = torch.tensor([10.0] * 3, device=device)
    cam = camera(dr, device, res, res)
    integ = dr.PathIntegrator(spp_interior=spp, max_bounce=3, spp_primary_edge=spp_edge,
                              spp_secondary_edge=spp_edge)

    def scene(vertices, tex):
        sphere = dr.Mesh(vertices, uv, faces, faces, dr.DiffuseMaterial(tex))
        ground = dr.Mesh(ground_v, None, ground_f, None, dr.DiffuseMaterial(ground_c))
        light = dr.AreaLight(light_v, light_f, emit)
        return dr.Scene([cam], [sphere, ground], [light])

    return dict(V0=torch.tensor(V0, device=device), Vp=torch.tensor(Vp, device=device),
                tex_opt=texture(1, tex_res, device), tex_target=texture(2, tex_res, device), scene=scene,
                integrator=integ, res=res, spp=spp, spp_edge=spp_edge, n_tris=int(F.shape[0]) + 4)


def fbm_value_noise(n: int, octaves: int = 6, amplitude: float = 1.0, seed: int = 5) -> np.ndarray:
    """(n, n) fBm of value noise: octave k samples a (2^(k+2) + 1)^2 lattice of
    uniform [-1, 1] values (seed) with smoothstep-weighted bilinear interpolation,
    weight amplitude / 2^k; fp32."""
    rng = np.random.default_rng(seed)
    u = np.linspace(0.0, 1.0, n)
    out = np.zeros((n, n))
    for k in range(octaves):
        cells = 2 ** (k + 2)
        lat = rng.uniform(-1.0, 1.0, (cells + 1, cells + 1))
        x = u * cells
        i = np.minimum(x.astype(np.int64), cells - 1)
        f = x - i
        w = f * f * (3.0 - 2.0 * f)
        a = lat[i][:, i] * (1 - w)[None, :] + lat[i][:, i + 1] * w[None, :]
        b = lat[i + 1][:, i] * (1 - w)[None, :] + lat[i + 1][:, i + 1] * w[None, :]
        out += (amplitude / 2 ** k) * (a * (1 - w)[:, None] + b * w[:, None])
    return out.astype(np.float32)


def terrain_grid(n: int, size: float = 8.0):
    """Base vertices (x, 0, z) of an n x n grid over [-size/2, size/2]^2, faces
    (2 per cell) and per-vertex uvs."""
    t = np.linspace(-size / 2, size / 2, n)
    X, Z = np.meshgrid(t, t, indexing="xy")
    V = np.stack([X.ravel(), np.zeros(n * n), Z.ravel()], 1).astype(np.float32)
    idx = np.arange(n * n, dtype=np.int64).reshape(n, n)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    F = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)]).astype(np.int32)
    g = np.linspace(0.0, 1.0, n)
    G0, G1 = np.meshgrid(g, g, indexing="xy")
    uv = np.stack([G0.ravel(), G1.ravel()], 1).astype(np.float32)
    return V, F, uv


def c4_parts(dr, device, n=1024, res=1024, spp=64, depth=4, tex_res=1024, spp_edge=1):
    """C4: fBm terrain of n x n heights with a texture, lit by two quad lights;
    `scene(heights, tex)` builds the Scene with vertices (x, height, z)."""
    V, F, uv = terrain_grid(n)
    heights = torch.tensor(fbm_value_noise(n), device=device).reshape(-1)
    base = torch.tensor(V, device=device)
    faces = torch.tensor(F, device=device)
    uvs = torch.tensor(uv, device=device)
    lights = []
    for x, e in ((-2.0, 12.0), (2.5, 8.0)):
        Vl, Fl = quad(1.5, 4.0, facing_down=True)
        Vl = Vl + np.float32([x, 0.0, 0.0])
        lights.append((torch.tensor(Vl, device=device), torch.tensor(Fl, device=device),
                       torch.tensor([e, e, e], device=device)))
    cam = dr.PerspectiveCamera(look_from=torch.tensor([0.0, 6.0, 7.0], device=device),
                               look_at=torch.tensor([0.0, 0.0, 0.0], device=device),
                               up=torch.tensor([0.0, 1.0, 0.0], device=device),
                               vfov=torch.tensor([math.pi / 4.0], device=device), height=res, width=res)
    integ = dr.PathIntegrator(spp_interior=spp, max_bounce=depth, spp_primary_edge=spp_edge,
                              spp_secondary_edge=spp_edge)

    def objects(h, tex):
        verts = torch.stack([base[:, 0], h, base[:, 2]], 1)
        mesh = dr.Mesh(verts, uvs, faces, faces, dr.DiffuseMaterial(tex))
        return [cam], [mesh], [dr.AreaLight(v, f, e) for v, f, e in lights]

    def scene(h, tex):
        return dr.Scene(*objects(h, tex))

    return dict(heights=heights, tex=texture(1, tex_res, device), scene=scene, objects=objects, integrator=integ,
                res=res, spp=spp, n_tris=int(F.shape[0]) + 4)


@dataclass
class BatchScene:
    """One scene of the batched configuration: parameters (requires_grad), the
    target image, the integrator, and `scene(vertices, tex)` building its Scene."""
    vertices: torch.Tensor
    tex: torch.Tensor
    target: Optional[torch.Tensor]
    integrator: object
    scene: object
    res: int
    spp: int
    n_tris: int


def c3_specs(n_scenes=64, seed=4, res_choices=(128, 256, 512, 1024), spp_choices=(4, 16, 64),
             log_tris=(3.0, 5.3)):
    """Per-scene draws of C3 (one generator, seed 4, in scene order): geodesic
    frequency k (20 k^2 ~ 10^U(log_tris)), scale, rotation, resolution, spp and
    the generator state for the radial perturbation."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n_scenes):
        n_target = 10.0 ** rng.uniform(*log_tris)
        k = max(2, int(round(math.sqrt(n_target / 20.0))))
        scale = float(rng.uniform(0.5, 2.0))
        Q = random_rotation(rng)
        res = int(rng.choice(res_choices))
        spp = int(rng.choice(spp_choices))
        specs.append(dict(k=k, n_tris=20 * k * k, scale=scale, rotation=Q, res=res, spp=spp,
                          noise_seed=int(rng.integers(0, 2 ** 31))))
    return specs


def c3_parts(dr, device, n_scenes=64, seed=4, res_choices=(128, 256, 512, 1024), spp_choices=(4, 16, 64),
             log_tris=(3.0, 5.3), tex_res=256, depth=3, spp_edge=1, sigma=0.02, targets=True, share=True):
    """C3: `n_scenes` scenes of uneven size (see the module docstring), their
    native handles sharing one set of per-render buffers (rendered in turn,
    the buffers grow to the largest render)."""
    from . import ops
    specs = c3_specs(n_scenes, seed, res_choices, spp_choices, log_tris)
    Vg, Fg = quad(10.0, -1.0)
    Vl, Fl = quad(2.0, 4.0, facing_down=True)
    ground_v, ground_f = torch.tensor(Vg, device=device), torch.tensor(Fg, device=device)
    ground_c = torch.tensor([0.5] * 3, device=device)
    light_v, light_f = torch.tensor(Vl, device=device), torch.tensor(Fl, device=device)
    emit = torch.tensor([10.0] * 3, device=device)
    out = []
    for si, sp in enumerate(specs):
        V, F = geodesic_sphere(sp["k"])
        res, spp = sp["res"], sp["spp"]
        r = 1.0 + sigma * np.random.default_rng(sp["noise_seed"]).standard_normal(V.shape[0])
        V0 = (V.astype(np.float64) @ sp["rotation"].T) * sp["scale"]
        uv = torch.tensor(sphere_uv(V), device=device)   # the texture moves with the sphere
        faces = torch.tensor(F, device=device)
        cam = camera(dr, device, res, res)
        integ = dr.PathIntegrator(spp_interior=spp, max_bounce=depth, spp_primary_edge=spp_edge,
                                  spp_secondary_edge=spp_edge)

        def scene(vertices, tex, uv=uv, faces=faces, cam=cam):
            sphere = dr.Mesh(vertices, uv, faces, faces, dr.DiffuseMaterial(tex))
            ground = dr.Mesh(ground_v, None, ground_f, None, dr.DiffuseMaterial(ground_c))
            light = dr.AreaLight(light_v, light_f, emit)
            return dr.Scene([cam], [sphere, ground], [light])

        v_opt = torch.tensor((V0 * r[:, None]).astype(np.float32), device=device)
        tex_opt = texture(1000 + 2 * si, tex_res, device)
        target = None
        if targets:
            with torch.no_grad():
                target = integ.renderC(scene(torch.tensor(V0.astype(np.float32), device=device),
                                             texture(1001 + 2 * si, tex_res, device)))[0]
        out.append(BatchScene(v_opt.requires_grad_(True), tex_opt.requires_grad_(True), target, integ, scene,
                              res, spp, int(F.shape[0]) + 4))
    if share:
        ops.share_workspace([b.scene(b.vertices, b.tex) for b in out])
    return out


def c3_step(parts, scene_step):
    """One step over every scene: scene_step(b) (render, loss, backward) for
    each scene in turn on the caller's stream."""
    for b in parts:
        scene_step(b)
