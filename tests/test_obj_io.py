"""OBJ / MTL loading (CPU): the examples' load_obj helper restated without the
external parser (example/optimize_texture.py:12-41)."""
import numpy as np
import torch

import torch_psdr_amd as dr
from torch_psdr_amd.flat import flatten

OBJ = """# two named meshes and a light
mtllib scene.mtl
v -1 0 -1
v 1 0 -1
v 1 0 1
v -1 0 1
v -0.5 2 -0.5
v 0.5 2 -0.5
v 0.5 2 0.5
v -0.5 2 0.5
v 0 0.5 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 1 0
o floor
usemtl grey
f 1/1/1 2/2/1 3/3/1 4/4/1
o light
usemtl lamp
f 8 7 6 5
o tri
usemtl tex
f -9/1 -8/2 -1/3
"""

MTL = """newmtl grey
Kd 0.25 0.5 0.75
newmtl lamp
Ka 7 8 9
Kd 0 0 0
newmtl tex
Kd 0.1 0.2 0.3
map_Kd checker.png
"""


def _write(tmp_path):
    from PIL import Image
    (tmp_path / "scene.obj").write_text(OBJ)
    (tmp_path / "scene.mtl").write_text(MTL)
    img = np.zeros((4, 4, 3), np.uint8)
    img[::2, ::2] = 255
    Image.fromarray(img).save(tmp_path / "checker.png")


def test_read_obj_faces_uvs_and_names(tmp_path):
    _write(tmp_path)
    d = dr.read_obj(tmp_path / "scene.obj")
    assert len(d.vertices) == 9 and len(d.tex_coords) == 4
    assert [m.name for m in d.meshes] == ["floor", "light", "tri"]
    floor, light, tri = d.meshes
    assert floor.faces == [(0, 1, 2), (0, 2, 3)] and floor.uv_faces == [(0, 1, 2), (0, 2, 3)]
    assert light.faces == [(7, 6, 5), (7, 5, 4)] and light.uv_faces == [(-1, -1, -1)] * 2
    assert tri.faces == [(0, 1, 8)] and tri.uv_faces == [(0, 1, 2)]
    mats = dr.read_mtl(tmp_path / "scene.mtl")
    assert mats["grey"].diffuse == (0.25, 0.5, 0.75) and mats["lamp"].ambient == (7.0, 8.0, 9.0)
    assert mats["tex"].diffuse_map == "checker.png"


def test_load_obj_builds_scene_objects(tmp_path, oracle):
    _write(tmp_path)
    objs = dr.load_obj(tmp_path / "scene.obj", device="cpu")
    assert isinstance(objs["light"], dr.AreaLight)
    np.testing.assert_array_equal(objs["light"].emit.numpy(), [7, 8, 9])
    floor, tri = objs["floor"], objs["tri"]
    assert isinstance(floor.material, dr.DiffuseBsdfMaterial)
    np.testing.assert_array_equal(floor.material.color.data.numpy(), [0.25, 0.5, 0.75])
    assert float(floor.material.roughness.data) == 0.5
    assert tri.material.color.tex_type == "UV3f" and tri.material.color.data.shape == (4, 4, 3)
    assert floor.vertices.shape == (9, 3) and floor.uv_indices is not None
    # the loaded objects flatten and render like the same scene built by hand
    flat = flatten([floor, tri], [objs["light"]], torch.device("cpu"))
    V = torch.tensor([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], dtype=torch.float32)
    quad = torch.tensor([[0, 1, 2], [0, 2, 3]], dtype=torch.int32)
    hand_floor = dr.Mesh(V, torch.tensor([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=torch.float32), quad, quad,
                         dr.DiffuseBsdfMaterial(torch.tensor([0.25, 0.5, 0.75]), torch.tensor([0.5])))
    L = torch.tensor([[-0.5, 2, -0.5], [0.5, 2, -0.5], [0.5, 2, 0.5], [-0.5, 2, 0.5]], dtype=torch.float32)
    hand_light = dr.AreaLight(L, torch.tensor([[3, 2, 1], [3, 1, 0]], dtype=torch.int32), torch.tensor([7., 8., 9.]))
    T = torch.tensor([[-1, 0, -1], [1, 0, -1], [0, 0.5, 0]], dtype=torch.float32)
    tex = torch.tensor([[[1.0] * 3, [0.0] * 3] * 2, [[0.0] * 3] * 4] * 2)
    hand_tri = dr.Mesh(T, torch.tensor([[0, 0], [1, 0], [1, 1]], dtype=torch.float32),
                       torch.tensor([[0, 1, 2]], dtype=torch.int32), torch.tensor([[0, 1, 2]], dtype=torch.int32),
                       dr.DiffuseBsdfMaterial(tex, torch.tensor([0.5])))
    np.testing.assert_array_equal(tri.material.color.data.numpy(), tex.numpy())
    hand = flatten([hand_floor, hand_tri], [hand_light], torch.device("cpu"))
    cp = [0.0, 1.0, 4.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.8, 12.0, 12.0]
    cfg = dr.PathIntegrator(spp_interior=2, max_bounce=2).cfg
    a = oracle.render(oracle.Scene(flat), cfg, cp)
    b = oracle.render(oracle.Scene(hand), cfg, cp)
    assert a.sum() > 0
    np.testing.assert_array_equal(a, b)
