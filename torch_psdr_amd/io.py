"""Wavefront OBJ / MTL loading into scene objects.

The reference's examples load scenes with the external `pywavefront_uv` package
and a small helper (`load_obj`, example/optimize_texture.py:12-41): every named
mesh of the file shares the file's vertex and texture-coordinate arrays and
gets its own triangle and uv indices; the mesh named "light" becomes an
AreaLight whose radiance is the material's ambient colour (Ka), every other
mesh a DiffuseBsdfMaterial with the material's diffuse colour (Kd, or its
map_Kd image) and roughness 0.5.  This module restates that helper without the
external parser: `read_obj` / `read_mtl` parse the formats, `load_obj` builds
the objects.

Supported OBJ statements: v, vt, f (v, v/vt, v/vt/vn, v//vn; negative indices;
polygons are split into triangle fans as pywavefront does), o / g (mesh name),
usemtl, mtllib.  MTL: newmtl, Kd, Ka, map_Kd (read with Pillow).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Tuple, Union

import torch

from . import api


@dataclass
class ObjMesh:
    name: Optional[str]
    material: Optional[str]
    faces: List[Tuple[int, int, int]] = field(default_factory=list)      # 0-based vertex ids
    uv_faces: List[Tuple[int, int, int]] = field(default_factory=list)   # 0-based tex-coord ids, -1: none


@dataclass
class ObjData:
    vertices: List[Tuple[float, float, float]]
    tex_coords: List[Tuple[float, float]]
    meshes: List[ObjMesh]
    mtllibs: List[str]


@dataclass
class MtlMaterial:
    diffuse: Tuple[float, float, float] = (0.8, 0.8, 0.8)
    ambient: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    diffuse_map: Optional[str] = None


def _index(tok: str, n: int) -> int:
    i = int(tok)
    return i - 1 if i > 0 else n + i


def read_obj(path: Union[str, Path]) -> ObjData:
    """Parse a Wavefront OBJ file (positions, texture coordinates, named meshes)."""
    verts: List[Tuple[float, float, float]] = []
    tcs: List[Tuple[float, float]] = []
    meshes: List[ObjMesh] = []
    mtllibs: List[str] = []
    name: Optional[str] = None
    mat: Optional[str] = None
    cur: Optional[ObjMesh] = None
    for raw in Path(path).read_text().splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        parts = line.split()
        key, args = parts[0], parts[1:]
        if key == "v":
            verts.append((float(args[0]), float(args[1]), float(args[2])))
        elif key == "vt":
            tcs.append((float(args[0]), float(args[1]) if len(args) > 1 else 0.0))
        elif key in ("o", "g"):
            name = " ".join(args) if args else None
            cur = None
        elif key == "usemtl":
            mat = " ".join(args)
            cur = None
        elif key == "mtllib":
            mtllibs.extend(args)
        elif key == "f":
            if cur is None:
                cur = ObjMesh(name, mat)
                meshes.append(cur)
            vi, ti = [], []
            for a in args:
                sub = a.split("/")
                vi.append(_index(sub[0], len(verts)))
                ti.append(_index(sub[1], len(tcs)) if len(sub) > 1 and sub[1] else -1)
            for k in range(1, len(vi) - 1):  # triangle fan
                cur.faces.append((vi[0], vi[k], vi[k + 1]))
                cur.uv_faces.append((ti[0], ti[k], ti[k + 1]))
    return ObjData(verts, tcs, [m for m in meshes if m.faces], mtllibs)


def read_mtl(path: Union[str, Path]) -> Dict[str, MtlMaterial]:
    """Parse a Wavefront MTL file (Kd, Ka, map_Kd per material)."""
    out: Dict[str, MtlMaterial] = {}
    cur: Optional[MtlMaterial] = None
    for raw in Path(path).read_text().splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        parts = line.split()
        key, args = parts[0], parts[1:]
        if key == "newmtl":
            cur = out.setdefault(" ".join(args), MtlMaterial())
        elif cur is None:
            continue
        elif key == "Kd":
            cur.diffuse = tuple(float(x) for x in args[:3])
        elif key == "Ka":
            cur.ambient = tuple(float(x) for x in args[:3])
        elif key == "map_Kd":
            cur.diffuse_map = args[-1]
    return out


def read_image(path: Union[str, Path]) -> torch.Tensor:
    """An RGB image as an (H, W, 3) float tensor in [0, 1] (row 0 = top)."""
    try:
        from PIL import Image
    except ImportError as e:  # pragma: no cover - Pillow is part of this image
        raise RuntimeError("reading texture images needs Pillow") from e
    import numpy as np
    img = np.asarray(Image.open(path).convert("RGB"), dtype=np.float32) / 255.0
    return torch.from_numpy(img.copy())


def load_obj(path: Union[str, Path], device: Union[str, torch.device] = "cuda",
             roughness: float = 0.5) -> Dict[str, Union[api.Mesh, api.AreaLight]]:
    """Scene objects of an OBJ file, named by mesh: the mesh named "light" as an
    AreaLight (radiance = Ka), the others as DiffuseBsdfMaterial meshes (Kd or
    map_Kd, the given roughness), all sharing the file's vertex and
    texture-coordinate arrays (example/optimize_texture.py:12-41)."""
    path = Path(path)
    data = read_obj(path)
    mats: Dict[str, MtlMaterial] = {}
    for lib in data.mtllibs:
        p = path.parent / lib
        if p.exists():
            mats.update(read_mtl(p))
    device = torch.device(device)
    vertices = torch.tensor(data.vertices, dtype=torch.float32, device=device)
    uvs = torch.tensor(data.tex_coords, dtype=torch.float32, device=device) if data.tex_coords else None
    objs: Dict[str, Union[api.Mesh, api.AreaLight]] = {}
    for m in data.meshes:
        name = m.name if m.name is not None else str(id(m))
        faces = torch.tensor(m.faces, dtype=torch.int32, device=device)
        mat = mats.get(m.material or "", MtlMaterial())
        if name == "light":
            objs[name] = api.AreaLight(vertices, faces, torch.tensor(mat.ambient, dtype=torch.float32, device=device))
            continue
        has_uv = uvs is not None and all(t >= 0 for f in m.uv_faces for t in f)
        uv_idx = torch.tensor(m.uv_faces, dtype=torch.int32, device=device) if has_uv else None
        if mat.diffuse_map is not None:
            color = read_image(path.parent / mat.diffuse_map).to(device)
        else:
            color = torch.tensor(mat.diffuse, dtype=torch.float32, device=device)
        material = api.DiffuseBsdfMaterial(color, torch.tensor([roughness], dtype=torch.float32, device=device))
        objs[name] = api.Mesh(vertices, uvs if has_uv else None, faces, uv_idx, material)
    return objs
