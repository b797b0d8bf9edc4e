"""Flatten a Scene's objects into the global arrays the native library consumes.

Object order is meshes first, then the meshes of area lights (Scene::Scene,
src/scene.cc:8-20); triangle ids are global in that order (the OptiX SBT offset
convention of src/scene_optix.cc:160-168).  Only gathers and concatenations
happen here (no arithmetic), and they are torch ops, so gradients of the global
vertex / texel / emission buffers flow back to the user's tensors.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import torch

from . import _native as N

# default corner uvs when a mesh has none (src/mesh.cc:31-36)
_DEFAULT_UV = (0.0, 0.0, 0.0, 1.0, 1.0, 0.0)

# The environment lights' bounding cube (EnvLight::set_aabb / SH9Light::set_aabb,
# src/emitter.cc:98-130, 218-247): corner k has x from bit 2, y from bit 1, z
# from bit 0 of k (0 = min, 1 = max); the faces and the cube map's uv layout
# (a 4 x 3 grid of unit squares, u step 1/4, v step 1/3) are the reference's,
# since the cube map texture is addressed through them.
_CUBE_FACES = ((0, 1, 2), (3, 2, 1), (4, 5, 6), (7, 6, 5), (0, 4, 1), (5, 1, 4),
               (2, 6, 3), (7, 3, 6), (4, 0, 6), (2, 6, 0), (5, 1, 7), (3, 7, 1))
_CUBE_UV_IDX = ((2, 3, 7), (8, 7, 3), (5, 4, 10), (9, 10, 4), (0, 1, 3), (4, 3, 1),
                (12, 13, 8), (9, 8, 13), (5, 6, 10), (11, 10, 6), (4, 3, 9), (8, 9, 3))


_CONST: dict = {}


def _const(device: torch.device, name: str, make) -> torch.Tensor:
    """A constant tensor on `device`, built on the host once and cached: a step
    that rebuilds its Scene then issues no host-to-device copy (such a copy
    synchronises the host, and cannot be recorded into a HIP graph)."""
    dev = torch.device(device)
    if dev.type == "cuda" and dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    key = (dev.type, dev.index, name)
    t = _CONST.get(key)
    if t is None:
        t = make().to(dev)
        _CONST[key] = t
    return t


def _default_uv(device) -> torch.Tensor:
    return _const(device, "default_uv", lambda: torch.tensor(_DEFAULT_UV, dtype=torch.float32))


def _dev_scalars(values, device) -> torch.Tensor:
    """A small fp32 vector on `device` made by fill kernels (no host copy)."""
    return torch.stack([torch.full((), float(x), dtype=torch.float32, device=device) for x in values])


def _cube_uvs():
    u, v = 0.25, 1.0 / 3.0
    f32 = lambda x: float(torch.tensor(x, dtype=torch.float32))
    u, v = f32(u), f32(v)
    us = (u, 2 * u, 0, u, 2 * u, 3 * u, 4 * u, 0, u, 2 * u, 3 * u, 4 * u, u, 2 * u)
    vs = (0, 0, v, v, v, v, v, 2 * v, 2 * v, 2 * v, 2 * v, 2 * v, 1, 1)
    return [(f32(a), f32(b)) for a, b in zip(us, vs)]


def _env_cube_vertices(cameras, meshes, device) -> torch.Tensor:
    """Scene::Scene's cube (src/scene.cc:31-50): the box of the camera positions
    and the meshes' vertex boxes, grown to its centre +- 1000 x its largest
    extent; all fp32 elementwise ops, detached."""
    with torch.no_grad():
        # camera positions from their host copies (BaseCamera.params)
        pos = [c.params()[:3] for c in cameras]
        lo3 = [min(p[a] for p in pos) for a in range(3)]
        hi3 = [max(p[a] for p in pos) for a in range(3)]
        pmin, pmax = _dev_scalars(lo3, device), _dev_scalars(hi3, device)
        for m in meshes:
            v = m.vertices.detach().to(torch.float32)
            pmin = torch.minimum(pmin, v.min(0).values)
            pmax = torch.maximum(pmax, v.max(0).values)
        d = (pmax - pmin).max() * 1000.0
        c = (pmin + pmax) * 0.5
        lo, hi = c - d, c + d
        corners = [torch.stack([(hi if (k >> 2) & 1 else lo)[0], (hi if (k >> 1) & 1 else lo)[1],
                                (hi if k & 1 else lo)[2]]) for k in range(8)]
        return torch.stack(corners)


@dataclass
class FlatScene:
    vertices: torch.Tensor          # (V, 3) f32, autograd-connected
    faces: torch.Tensor             # (F, 3) i32 global vertex ids
    tri_uv: torch.Tensor            # (F, 6) f32
    tri_obj: torch.Tensor           # (F,) i32
    texels: torch.Tensor            # (T,) f32, autograd-connected
    light_emit: torch.Tensor        # (L, 3) f32, autograd-connected
    objects: List[dict]             # psdr_object fields per object
    light_obj: List[int]
    obj_vertex_range: List[tuple] = field(default_factory=list)
    topo_key: tuple = ()

    @property
    def n_tris(self) -> int:
        return int(self.faces.shape[0])

    def desc(self, keep_alive: list) -> N.PsdrSceneDesc:
        """ctypes psdr_scene_desc over this scene's device buffers."""
        objs = (N.PsdrObject * max(len(self.objects), 1))()
        for i, o in enumerate(self.objects):
            for k, v in o.items():
                setattr(objs[i], k, int(v))
        lobj = (N.C.c_int32 * max(len(self.light_obj), 1))(*self.light_obj)
        keep_alive += [objs, lobj]
        d = N.PsdrSceneDesc()
        d.n_vertices = int(self.vertices.shape[0])
        d.vertices = self.vertices.data_ptr()
        d.n_tris = self.n_tris
        d.faces = self.faces.data_ptr()
        d.tri_uv = self.tri_uv.data_ptr()
        d.tri_obj = self.tri_obj.data_ptr()
        d.n_objects = len(self.objects)
        d.objects = N.C.cast(objs, N.C.POINTER(N.PsdrObject))
        d.n_texels = int(self.texels.numel())
        d.texels = self.texels.data_ptr()
        d.n_lights = len(self.light_obj)
        d.light_emit = self.light_emit.data_ptr() if self.light_emit.numel() else None
        d.light_obj = N.C.cast(lobj, N.C.POINTER(N.C.c_int32))
        return d


def _tex_entry(tex, pieces: list, offset: int):
    """Append a texture's data; returns (offset, h, w, new_offset)."""
    data = tex.data
    if tex.is_constant:
        flat = data.reshape(-1)
        pieces.append(flat)
        return offset, 0, 0, offset + flat.numel()
    flat = data.reshape(-1)
    pieces.append(flat)
    return offset, tex.height, tex.width, offset + flat.numel()


def flatten(meshes, lights, device, env=None, cameras=()) -> FlatScene:
    """meshes: list of api.Mesh; lights: list of api.AreaLight; env: an
    api.EnvLight / api.SH9Light (its bounding cube becomes the last object)."""
    objs = [(m, None) for m in meshes] + [(l.mesh_proxy(), l) for l in lights]
    verts, faces, uvs, tobj = [], [], [], []
    objects, light_obj = [], []
    tex_pieces: list = []
    tex_off = 0
    voff = 0
    toff = 0
    vranges = []
    key = []
    emits = []
    for oi, (m, light) in enumerate(objs):
        v = m.vertices
        f = m.face_indices
        if v.device != device or f.device != device:
            raise RuntimeError("all scene tensors must live on the same GPU device")
        nv, nf = int(v.shape[0]), int(f.shape[0])
        verts.append(v.to(torch.float32))
        faces.append(f.to(torch.int32) + voff)
        if m.uvs is not None and m.uv_indices is not None:
            corner = m.uvs.to(torch.float32)[m.uv_indices.long()]  # (F, 3, 2)
            uvs.append(corner.reshape(nf, 6))
        else:
            uvs.append(_default_uv(device).expand(nf, 6))
        tobj.append(torch.full((nf,), oi, dtype=torch.int32, device=device))
        rec = dict(material=N.PSDR_MAT_DIFFUSE, light=-1, albedo_offset=0, albedo_h=0, albedo_w=0,
                   rough_offset=0, rough_h=0, rough_w=0, tri_begin=toff, tri_count=nf, grad_mask=0,
                   normal_offset=-1, normal_h=0, normal_w=0)
        if v.requires_grad:
            rec["grad_mask"] |= N.PSDR_GRAD_VERTICES
        if light is not None:
            rec["material"] = N.PSDR_MAT_LIGHT
            rec["light"] = len(light_obj)
            light_obj.append(oi)
            emits.append(light.emit.to(torch.float32).reshape(3))
            if light.emit.requires_grad:
                rec["grad_mask"] |= N.PSDR_GRAD_EMIT
        else:
            mat = m.material
            a_off, a_h, a_w, tex_off = _tex_entry(mat.color, tex_pieces, tex_off)
            rec.update(albedo_offset=a_off, albedo_h=a_h, albedo_w=a_w)
            if mat.color.data.requires_grad:
                rec["grad_mask"] |= N.PSDR_GRAD_ALBEDO
            if mat.kind == "diffuse_bsdf":
                rec["material"] = N.PSDR_MAT_DIFFUSE_BSDF
                r_off, r_h, r_w, tex_off = _tex_entry(mat.roughness, tex_pieces, tex_off)
                rec.update(rough_offset=r_off, rough_h=r_h, rough_w=r_w)
                if mat.roughness.data.requires_grad:
                    rec["grad_mask"] |= N.PSDR_GRAD_ROUGHNESS
                if mat.normal is not None:
                    n_off, n_h, n_w, tex_off = _tex_entry(mat.normal, tex_pieces, tex_off)
                    rec.update(normal_offset=n_off, normal_h=n_h, normal_w=n_w)
                    if mat.normal.data.requires_grad:
                        rec["grad_mask"] |= N.PSDR_GRAD_NORMAL
        objects.append(rec)
        vranges.append((voff, voff + nv))
        key.append((f.data_ptr(), f._version, tuple(f.shape), nv,
                    None if m.uv_indices is None else (m.uv_indices.data_ptr(), m.uv_indices._version),
                    None if m.uvs is None else (m.uvs.data_ptr(), m.uvs._version)))
        voff += nv
        toff += nf
    if env is not None:
        oi = len(objs)
        verts.append(_env_cube_vertices(cameras, meshes, device))
        faces.append(_const(device, "cube_faces", lambda: torch.tensor(_CUBE_FACES, dtype=torch.int32)) + voff)
        rec = dict(light=len(light_obj), rough_offset=0, rough_h=0, rough_w=0, tri_begin=toff, tri_count=12,
                   grad_mask=0, normal_offset=-1, normal_h=0, normal_w=0)
        if hasattr(env, "cube_map"):
            uvs.append(_const(device, "cube_uv", lambda: torch.tensor(_cube_uvs(), dtype=torch.float32)[
                torch.tensor(_CUBE_UV_IDX).long()].reshape(12, 6)))
            a_off, a_h, a_w, tex_off = _tex_entry(env.cube_map, tex_pieces, tex_off)
            rec.update(material=N.PSDR_MAT_ENV_MAP, albedo_offset=a_off, albedo_h=a_h, albedo_w=a_w)
            if env.cube_map.data.requires_grad:
                rec["grad_mask"] |= N.PSDR_GRAD_ALBEDO
            key.append(("env_map",))
        else:
            uvs.append(_default_uv(device).expand(12, 6))
            coeff = env.sh9_coeff.to(device=device, dtype=torch.float32).reshape(27)
            tex_pieces.append(coeff)
            rec.update(material=N.PSDR_MAT_ENV_SH9, albedo_offset=tex_off, albedo_h=0, albedo_w=0)
            tex_off += 27
            if env.sh9_coeff.requires_grad:
                rec["grad_mask"] |= N.PSDR_GRAD_ALBEDO
            key.append(("env_sh9",))
        tobj.append(torch.full((12,), oi, dtype=torch.int32, device=device))
        light_obj.append(oi)
        emits.append(torch.zeros(3, dtype=torch.float32, device=device))
        objects.append(rec)
        vranges.append((voff, voff + 8))
        voff += 8
        toff += 12
    texels = torch.cat(tex_pieces) if tex_pieces else torch.zeros(1, dtype=torch.float32, device=device)
    emit = torch.stack(emits) if emits else torch.zeros((0, 3), dtype=torch.float32, device=device)
    return FlatScene(
        vertices=torch.cat(verts).contiguous() if verts else torch.zeros((0, 3), device=device),
        faces=torch.cat(faces).contiguous() if faces else torch.zeros((0, 3), dtype=torch.int32, device=device),
        tri_uv=torch.cat(uvs).contiguous() if uvs else torch.zeros((0, 6), device=device),
        tri_obj=torch.cat(tobj).contiguous() if tobj else torch.zeros((0,), dtype=torch.int32, device=device),
        texels=texels.contiguous(),
        light_emit=emit.contiguous(),
        objects=objects,
        light_obj=light_obj,
        obj_vertex_range=vranges,
        topo_key=tuple(key),
    )
