"""Scene builders shared by the tests (CPU or GPU tensors)."""
from __future__ import annotations

import math

import numpy as np
import torch

import torch_psdr_amd as dr
from torch_psdr_amd import workloads as W
from torch_psdr_amd.flat import flatten


def objects(device, level=1, tex_res=None, tex_seed=1, sphere_vertices=None, light_y=4.0, emit=10.0,
            bsdf_rough=None, normal_res=None, requires_grad=False, env=None, env_res=6, area_light=True,
            ground=True):
    """Sphere (+ optional UV texture, DiffuseBsdf roughness, normal map), ground quad, quad
    light, optional environment light (env = "map": a random env_res x 4 env_res/3 cube map,
    "sh9": random SH coefficients): (meshes, lights, leaves)."""
    Vs, Fs = W.icosphere(level)
    if sphere_vertices is not None:
        Vs = sphere_vertices
    vs = torch.tensor(Vs, device=device, requires_grad=requires_grad)
    fs = torch.tensor(Fs, device=device)
    leaves = {"sphere_v": vs}
    if tex_res is None:
        color = torch.tensor([0.5, 0.5, 0.5], device=device, requires_grad=requires_grad)
        uv, uvi = None, None
        if normal_res is not None:
            uv = torch.tensor(W.sphere_uv(W.icosphere(level)[0]), device=device)
            uvi = fs
    else:
        g = torch.Generator().manual_seed(tex_seed)
        color = (0.2 + 0.6 * torch.rand((tex_res, tex_res, 3), generator=g)).to(device)
        color.requires_grad_(requires_grad)
        uv = torch.tensor(W.sphere_uv(W.icosphere(level)[0]), device=device)
        uvi = fs
    leaves["color"] = color
    if bsdf_rough is None:
        mat = dr.DiffuseMaterial(color)
    else:
        rough = torch.tensor([bsdf_rough], device=device, requires_grad=requires_grad)
        leaves["rough"] = rough
        normal = None
        if normal_res is not None:
            # tangent-space normals near +z: 2n-1 in [-0.3, 0.3]^2 x [0.6, 1]
            gn = torch.Generator().manual_seed(tex_seed + 100)
            r = torch.rand((normal_res, normal_res, 3), generator=gn)
            normal = torch.stack([0.35 + 0.3 * r[..., 0], 0.35 + 0.3 * r[..., 1], 0.8 + 0.2 * r[..., 2]], -1)
            normal = normal.to(device).requires_grad_(requires_grad)
            leaves["normal"] = normal
        mat = dr.DiffuseBsdfMaterial(color, rough, normal)
    sphere = dr.Mesh(vs, uv, fs, uvi, mat)
    meshes, lights = [sphere], []
    if ground:
        Vg, Fg = W.quad(10.0, -1.0)
        gv = torch.tensor(Vg, device=device, requires_grad=requires_grad)
        ground_col = torch.tensor([0.5, 0.5, 0.5], device=device, requires_grad=requires_grad)
        meshes.append(dr.Mesh(gv, None, torch.tensor(Fg, device=device), None, dr.DiffuseMaterial(ground_col)))
        leaves.update(ground_v=gv, ground_color=ground_col)
    if area_light:
        Vl, Fl = W.quad(2.0, light_y, facing_down=True)
        lv = torch.tensor(Vl, device=device, requires_grad=requires_grad)
        le = torch.tensor([emit, emit, emit], device=device, requires_grad=requires_grad)
        lights.append(dr.AreaLight(lv, torch.tensor(Fl, device=device), le))
        leaves.update(light_v=lv, emit=le)
    if env is not None:
        ge = torch.Generator().manual_seed(tex_seed + 200)
        if env == "map":
            cube = (0.5 + 1.5 * torch.rand((env_res, 4 * env_res // 3 + 1, 3), generator=ge)).to(device)
            cube.requires_grad_(requires_grad)
            leaves["env_map"] = cube
            lights.append(dr.EnvLight(dr.Texture(cube)))
        else:
            coeff = torch.rand((9, 3), generator=ge) * 0.6 - 0.2
            coeff[0] = 2.0 + torch.rand(3, generator=ge)
            coeff = coeff.to(device).requires_grad_(requires_grad)
            leaves["sh9"] = coeff
            lights.append(dr.SH9Light(coeff))
    return meshes, lights, leaves


def cam_params(res, look_from=(0.0, 1.0, 5.0), vfov=math.pi / 4):
    return list(look_from) + [0.0, 0.0, 0.0, 0.0, 1.0, 0.0, vfov, float(res), float(res)]


def camera(device, res, look_from=(0.0, 1.0, 5.0), vfov=math.pi / 4, requires_grad=False):
    t = lambda v: torch.tensor(v, device=device, requires_grad=requires_grad)
    return dr.PerspectiveCamera(t(list(look_from)), t([0.0, 0.0, 0.0]), t([0.0, 1.0, 0.0]), t([vfov]), res, res)


# orthographic camera of the same view: a 4.5-wide window looking at the origin
def ortho_params(res, look_from=(0.0, 1.0, 5.0), view_width=4.5):
    return list(look_from) + [0.0, 0.0, 0.0, 0.0, 1.0, 0.0, view_width, float(res), float(res), 1.0]


def ortho_camera(device, res, look_from=(0.0, 1.0, 5.0), view_width=4.5, requires_grad=False):
    t = lambda v: torch.tensor(v, device=device, requires_grad=requires_grad)
    return dr.OrthographicCamera(t(list(look_from)), t([0.0, 0.0, 0.0]), t([0.0, 1.0, 0.0]), t([view_width]), res,
                                 res)


def cpu_flat(**kw):
    meshes, lights, leaves = objects(torch.device("cpu"), **kw)
    area = [l for l in lights if isinstance(l, dr.AreaLight)]
    env = [l for l in lights if not isinstance(l, dr.AreaLight)]
    cams = [camera(torch.device("cpu"), 8)]
    return flatten(meshes, area, torch.device("cpu"), env=env[-1] if env else None, cameras=cams), leaves


def tri9(flat) -> np.ndarray:
    V = flat.vertices.detach().cpu().numpy()
    F = flat.faces.cpu().numpy()
    return V[F].reshape(-1, 9).astype(np.float32)
