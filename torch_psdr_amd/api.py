"""Drop-in Python API of torch_psdr (src/python/torch_psdr.py) on the gfx950 library.

Same class names, constructor arguments, shapes, dtypes, defaults and
assertions as the reference wrapper; `PathIntegrator.renderC` / `renderD`
return a list of (H, W, 3) fp32 images, one per camera (row 0 = top), and
`renderD`'s images are differentiable w.r.t. mesh / light vertices, albedo,
roughness and normal-map texels, light emission and camera parameters through
torch autograd.

Differences, each deliberate:
  * 2-D (H, W) textures are accepted as 1-channel maps (the reference's `==`
    typo at src/python/torch_psdr.py:233 rejects them);
  * the stray `from turtle import forward` (src/python/torch_psdr.py:2) is gone.
"""
from __future__ import annotations

import warnings
from typing import List, Union

import torch
from torch import Tensor

from . import ops
from .flat import flatten


class VarContex:
    """Bookkeeping of tensors that require grad (src/python/torch_psdr.py:84-103)."""

    def __init__(self):
        self.torch_vars: List[Tensor] = []

    def append_variable(self, torch_var: Tensor, native_var=None):
        if torch_var.requires_grad:
            self.torch_vars.append(torch_var)

    def get_variables(self):
        return [], self.torch_vars

    def merge_variables(self, native_vars, torch_vars):
        for t in torch_vars:
            if t.requires_grad:
                self.torch_vars.append(t)


def _capturing() -> bool:
    return torch.cuda.is_available() and torch.cuda.is_current_stream_capturing()


class BaseCamera(VarContex):
    """Base camera class (src/python/torch_psdr.py:106-110).

    The parameters (look_from, look_at, up and vfov or view_width) may require
    grad: their gradient flows, as in the reference, only through ray origins
    (the camera vertex's direction to the viewer and the primary boundary's
    origin term)."""

    kind = 0

    def _init(self, look_from: Tensor, look_at: Tensor, up: Tensor, extent: Tensor, height: int, width: int):
        super().__init__()
        assert look_from.size() == (3,)
        assert look_at.size() == (3,)
        assert up.size() == (3,)
        assert extent.size() == (1,)
        assert height > 0 and width > 0
        for t in (look_from, look_at, up, extent):
            self.append_variable(t)
        self.look_from, self.look_at, self.up = look_from, look_at, up
        self.height = int(height)
        self.width = int(width)
        self._extent = extent

    def param_tensor(self) -> Tensor:
        """The PSDR_CAMERA_GRAD parameters as one (differentiable) fp32 vector."""
        parts = [self.look_from.reshape(3), self.look_at.reshape(3), self.up.reshape(3), self._extent.reshape(1)]
        dev = next((t.device for t in parts if t.device.type != "cpu"), parts[0].device)
        return torch.cat([t.to(device=dev, dtype=torch.float32) for t in parts])

    def params(self) -> List[float]:
        """Host values of the parameters (+ height, width, kind).  Read from the
        tensors at every eager call; under HIP graph capture, where a copy from
        device tensors is not possible, the values of the last eager read are
        used, provided the tensors have not changed since."""
        parts = (self.look_from, self.look_at, self.up, self._extent)
        key = tuple((t.data_ptr(), t._version) for t in parts)
        if _capturing() and any(t.is_cuda for t in parts):
            if getattr(self, "_host_key", None) != key:
                raise RuntimeError("camera parameters changed (or were never read) since the last eager render: a "
                                   "captured step renders with camera values read before capture, so render once "
                                   "eagerly first, and keep camera optimisation in eager steps")
            vals = self._host_vals
        else:
            vals = self.param_tensor().detach().cpu().tolist()
            self._host_vals, self._host_key = vals, key
        return vals + [float(self.height), float(self.width), float(self.kind)]


class PerspectiveCamera(BaseCamera):
    """PerspectiveCamera(look_from (3,), look_at (3,), up (3,), vfov (1,) [rad], height, width)
    (src/python/torch_psdr.py:113-158; matrices src/camera.h:7-53, near -0.01, far -1000)."""

    kind = 0

    def __init__(self, look_from: Tensor, look_at: Tensor, up: Tensor, vfov: Tensor, height: int, width: int):
        self._init(look_from, look_at, up, vfov, height, width)
        self.vfov = vfov


class OrthographicCamera(BaseCamera):
    """OrthographicCamera(look_from (3,), look_at (3,), up (3,), view_width (1,), height, width)
    (src/python/torch_psdr.py:161-206; projection src/camera.h:279-362): parallel
    rays along the view direction from a view_width-wide window on the camera plane."""

    kind = 1

    def __init__(self, look_from: Tensor, look_at: Tensor, up: Tensor, view_width: Tensor, height: int, width: int):
        self._init(look_from, look_at, up, view_width, height, width)
        self.view_width = view_width


class Texture(VarContex):
    """Constant or UV texture (src/python/torch_psdr.py:209-269).

    Shapes: (1,) / (3,) constant, (H, W) / (H, W, 1) / (H, W, 3) maps.  The data
    is kept in the caller's layout (row 0 = image top); lookups flip rows exactly
    as the reference's `tex.flip(0)` (nearest texel, src/texture.h:65-74)."""

    def __init__(self, tex: Tensor):
        super().__init__()
        tex_type = None
        dims = tex.dim()
        if dims == 1:
            if tex.size() == (1,):
                tex_type = "Constant1f"
            elif tex.size() == (3,):
                tex_type = "Constant3f"
        elif dims == 2:
            tex_type = "UV1f"
            tex = tex.unsqueeze(-1)
        elif dims == 3:
            if tex.size(-1) == 1:
                tex_type = "UV1f"
            elif tex.size(-1) == 3:
                tex_type = "UV3f"
        assert tex_type is not None
        self.tex_type = tex_type
        self.data = tex if tex.dtype == torch.float32 else tex.float()
        self.is_constant = tex_type.startswith("Constant")
        self.height = 0 if self.is_constant else int(tex.size(0))
        self.width = 0 if self.is_constant else int(tex.size(1))
        self.append_variable(tex)


class Material(VarContex):
    """Base class of materials (src/python/torch_psdr.py:272-277)."""


class DiffuseMaterial(Material):
    """Lambertian albedo texture (src/python/torch_psdr.py:280-295; src/diffuse_material.cc)."""

    kind = "diffuse"

    def __init__(self, tex: Union[Tensor, Texture]):
        super().__init__()
        if isinstance(tex, Tensor):
            tex = Texture(tex)
        assert tex.tex_type == "Constant3f" or tex.tex_type == "UV3f"
        self.color = tex
        self.merge_variables(*tex.get_variables())


class DiffuseBsdfMaterial(Material):
    """Lambertian / Oren-Nayar-style diffuse (src/python/torch_psdr.py:298-335; src/diffuse_bsdf.cc)."""

    kind = "diffuse_bsdf"

    def __init__(self, color: Union[Tensor, Texture], roughness: Union[Tensor, Texture],
                 normal: Union[Tensor, Texture, None] = None):
        super().__init__()
        if isinstance(color, Tensor):
            color = Texture(color)
        if isinstance(roughness, Tensor):
            roughness = Texture(roughness)
        if isinstance(normal, Tensor):
            normal = Texture(normal)
        assert color.tex_type == "Constant3f" or color.tex_type == "UV3f"
        assert roughness.tex_type == "Constant1f" or roughness.tex_type == "UV1f"
        assert normal is None or normal.tex_type == "Constant3f" or normal.tex_type == "UV3f"
        self.color = color
        self.roughness = roughness
        self.normal = normal
        self.merge_variables(*color.get_variables())
        self.merge_variables(*roughness.get_variables())
        if normal is not None:
            self.merge_variables(*normal.get_variables())


class PrincipledMaterial(Material):
    """Placeholder kept for name compatibility (empty in the reference too)."""


class Mesh(VarContex):
    """Triangle mesh + material (src/python/torch_psdr.py:343-379; src/mesh.cc:5-95)."""

    def __init__(self, vertices: Tensor, uvs: Union[None, Tensor], face_indices: Tensor,
                 uv_indices: Union[None, Tensor], material: Material):
        super().__init__()
        assert vertices.dim() == 2 and vertices.size(-1) == 3
        assert uvs is None or (uvs.dim() == 2 and uvs.size(-1) == 2)
        assert face_indices.dim() == 2 and face_indices.size(-1) == 3
        assert uv_indices is None or (uv_indices.dim() == 2 and uv_indices.size(-1) == 3)
        no_uvs = uvs is None or uv_indices is None
        self.vertices = vertices
        self.uvs = None if no_uvs else uvs
        self.face_indices = face_indices
        self.uv_indices = None if no_uvs else uv_indices
        self.material = material
        self.append_variable(vertices)
        self.merge_variables(*material.get_variables())


class Light(VarContex):
    """Base class of lights (src/python/torch_psdr.py:382-385)."""


class _LightMesh:
    def __init__(self, vertices, face_indices):
        self.vertices = vertices
        self.face_indices = face_indices
        self.uvs = None
        self.uv_indices = None
        self.material = None


class AreaLight(Light):
    """Triangle-mesh light with constant radiance (src/python/torch_psdr.py:388-414; src/emitter.cc:4-92)."""

    def __init__(self, vertices: Tensor, face_indices: Tensor, emit: Tensor):
        super().__init__()
        assert vertices.dim() == 2 and vertices.size(-1) == 3
        assert face_indices.dim() == 2 and face_indices.size(-1) == 3
        assert emit.dim() == 1 and emit.size(0) == 3
        self.vertices = vertices
        self.face_indices = face_indices
        self.emit = emit
        self.append_variable(vertices)
        self.append_variable(emit)

    def mesh_proxy(self):
        return _LightMesh(self.vertices, self.face_indices)


class EnvLight(Light):
    """Cube-map environment light (src/python/torch_psdr.py:417-430; src/emitter.cc:95-213): the
    radiance of a bounding cube around the cameras and meshes, looked up in `cube_map` (a UV3f
    texture laid out as the cube's unfolded faces) at the hit's uv."""

    def __init__(self, cube_map: Texture):
        super().__init__()
        assert cube_map.tex_type == "UV3f"
        self.cube_map = cube_map
        self.merge_variables(*cube_map.get_variables())


class SH9Light(Light):
    """Second-order spherical-harmonics environment light (src/python/torch_psdr.py:433-447;
    src/emitter.cc:215-352): radiance sum_k C_k Y_k(w) of the direction w from the bounding cube
    back to the shaded point, with the (9, 3) RGB coefficients C."""

    def __init__(self, sh9_coeff: Tensor):
        super().__init__()
        assert sh9_coeff.size() == (9, 3)
        self.sh9_coeff = sh9_coeff
        self.append_variable(sh9_coeff)


class Scene(VarContex):
    """Cameras, meshes and lights (src/python/torch_psdr.py:450-474; src/scene.cc:3-104).

    The scene's buffers are captured at construction (as the reference copies
    tensors into Enoki arrays), flattened into global vertex / texel / emission
    buffers on the device."""

    def __init__(self, cameras: List[BaseCamera], meshes: List[Mesh], lights: List[Light]):
        super().__init__()
        for c in cameras:
            self.merge_variables(*c.get_variables())
        for m in meshes:
            self.merge_variables(*m.get_variables())
        for l in lights:
            self.merge_variables(*l.get_variables())
        area = [l for l in lights if isinstance(l, AreaLight)]
        env = [l for l in lights if isinstance(l, (EnvLight, SH9Light))]
        if len(area) + len(env) != len(lights):
            raise TypeError("lights must be AreaLight, EnvLight or SH9Light")
        if len(env) > 1:
            # the reference reports this and keeps the last one (src/scene.cc:22-27)
            warnings.warn("only support no more than one env_light; using the last one")
        if not meshes and not area:
            raise ValueError("Scene needs at least one mesh or area light")
        if env and not cameras:
            raise ValueError("an environment light needs a camera (its cube is placed around the cameras)")
        device = (meshes[0].vertices if meshes else area[0].vertices).device
        if device.type != "cuda":
            raise RuntimeError("torch_psdr_amd renders on the GPU: scene tensors must be on a HIP device")
        self.cameras = cameras
        self.meshes = meshes
        self.lights = lights
        self.flat = flatten(meshes, area, device, env=env[-1] if env else None, cameras=cameras)
        self.token = ops.register(self.flat)

    def __del__(self):
        try:
            ops.unregister(self.token)
        except Exception:
            pass


class PathIntegrator:
    """Path-space differentiable integrator (src/python/torch_psdr.py:477-556; src/path_integrator.cc)."""

    def __init__(self, enable_light_visable: bool = False, n_pass: int = 1, spp_interior: int = 1,
                 max_bounce: int = 3, mis_light_samples: int = 1, mis_bsdf_samples: int = 1,
                 spp_primary_edge: int = 1, spp_secondary_edge: int = 1, primary_edge_preprocess_rounds: int = 1,
                 secondary_edge_preprocess_rounds: int = 1, primary_edge_hypercube_resolution: int = 10000,
                 secondary_edge_hypercube_resolution: List[int] = [10000, 6, 6]):
        assert n_pass > 0
        assert spp_interior > 0
        assert spp_primary_edge > 0
        assert spp_secondary_edge > 0
        assert max_bounce > 0
        assert mis_light_samples >= 0
        assert mis_bsdf_samples >= 0
        assert primary_edge_preprocess_rounds > 0
        assert secondary_edge_preprocess_rounds > 0
        assert primary_edge_hypercube_resolution > 0
        assert len(secondary_edge_hypercube_resolution) == 3
        for reso in secondary_edge_hypercube_resolution:
            assert reso > 0
        self.cfg = [int(n_pass), int(spp_interior), int(bool(enable_light_visable)), int(spp_primary_edge),
                    int(spp_secondary_edge), int(max_bounce), int(mis_light_samples), int(mis_bsdf_samples),
                    int(primary_edge_preprocess_rounds), int(secondary_edge_preprocess_rounds),
                    int(primary_edge_hypercube_resolution)] + [int(r) for r in secondary_edge_hypercube_resolution]

    # cameras per wavefront (the reference renders them one after another,
    # src/path_integrator.cc:385-403; here up to PSDR_MAX_CAMERAS share every launch)
    _CAMS_PER_LAUNCH = 4

    def _groups(self, scene: Scene):
        cams = scene.cameras
        return [cams[i:i + self._CAMS_PER_LAUNCH] for i in range(0, len(cams), self._CAMS_PER_LAUNCH)]

    def renderC(self, scene: Scene) -> List[Tensor]:
        ns, flat = ops.entry(scene.token)
        out: List[Tensor] = []
        with torch.no_grad():
            for group in self._groups(scene):
                if len(group) == 1:
                    out.append(ops.forward_native(ns, flat, group[0].params(), self.cfg, False))
                else:
                    out.extend(ops.forward_native_multi(ns, flat, [c.params() for c in group], self.cfg, False))
        return out

    def renderD(self, scene: Scene) -> List[Tensor]:
        f = scene.flat
        out: List[Tensor] = []
        for group in self._groups(scene):
            if len(group) == 1:
                cam = group[0]
                out.append(torch.ops.psdr.render(scene.token, cam.params(), self.cfg, f.vertices, f.texels,
                                                 f.light_emit, cam.param_tensor(), True))
            else:
                out.extend(torch.ops.psdr.render_multi(scene.token, ops.pack_cams([c.params() for c in group]),
                                                       self.cfg,
                                                       f.vertices, f.texels, f.light_emit,
                                                       [c.param_tensor() for c in group], True))
        return out
