"""Environment-light pins (CPU): the oracle's SH9 radiance against the defining
properties of the real spherical harmonics, the bounding cube against
Scene::Scene's closed form (src/scene.cc:31-55), and a uniform environment on a
convex object against the closed-form reflected radiance albedo * c."""
import math

import numpy as np
import pytest
import torch

import torch_psdr_amd as dr
from torch_psdr_amd import workloads as W
from torch_psdr_amd.flat import flatten
from tests import helpers as H

CPU = torch.device("cpu")


def _fibonacci_sphere(n):
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - z * z)
    phi = math.pi * (1.0 + math.sqrt(5.0)) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def test_sh9_basis_is_orthonormal(oracle):
    w = _fibonacci_sphere(60000)
    Y = np.zeros((9, w.shape[0]))
    for k in range(9):
        c = np.zeros((9, 3), np.float32)
        c[k, 0] = 1.0
        Y[k] = oracle.sh9_eval(c, w)[:, 0]
    G = (4.0 * math.pi / w.shape[0]) * Y @ Y.T
    np.testing.assert_allclose(G, np.eye(9), atol=2e-3)
    # the reference's order: Y1 ~ y, Y2 ~ z, Y3 ~ x (src/emitter.cc:258-260)
    for k, axis in ((1, 1), (2, 2), (3, 0)):
        e = np.zeros((1, 3))
        e[0, axis] = 1.0
        c = np.zeros((9, 3), np.float32)
        c[k, 0] = 1.0
        assert oracle.sh9_eval(c, e)[0, 0] == pytest.approx(math.sqrt(3.0 / (4.0 * math.pi)), rel=1e-6)
    # each channel is linear in its own coefficient column
    rng = np.random.default_rng(1)
    c = rng.normal(size=(9, 3)).astype(np.float32)
    got = oracle.sh9_eval(c, w[:50])
    np.testing.assert_allclose(got, (Y[:, :50].T @ c.astype(np.float64)), rtol=1e-5, atol=1e-5)


def test_environment_cube_closed_form():
    meshes, lights, _ = H.objects(CPU, level=1, env="map", area_light=False)
    cam = H.camera(CPU, 8, look_from=(0.5, 2.0, 6.0))
    flat = flatten(meshes, [], CPU, env=lights[-1], cameras=[cam])
    V = flat.vertices.numpy()[-8:]
    pts = np.concatenate([m.vertices.numpy() for m in meshes] + [np.float32([[0.5, 2.0, 6.0]])])
    lo, hi = pts.min(0), pts.max(0)
    d = np.float32(1000.0) * np.float32((hi - lo).max())
    c = (lo + hi) * np.float32(0.5)
    for k in range(8):
        want = [(c + d)[a] if (k >> (2 - a)) & 1 else (c - d)[a] for a in range(3)]
        np.testing.assert_array_equal(V[k], np.float32(want))
    ob = flat.objects[-1]
    assert ob["material"] == 3 and ob["tri_count"] == 12 and ob["light"] == len(flat.light_obj) - 1
    F = flat.faces.numpy()[-12:] - (flat.vertices.shape[0] - 8)
    # 12 triangles covering the 6 faces, each face split in two
    assert sorted(map(tuple, np.sort(F, 1).tolist())) == sorted(set(map(tuple, np.sort(F, 1).tolist())))
    face_planes = set()
    for f in F:
        corners = np.array([[(k >> 2) & 1, (k >> 1) & 1, k & 1] for k in f])
        const = [a for a in range(3) if (corners[:, a] == corners[0, a]).all()]
        assert len(const) == 1
        face_planes.add((const[0], int(corners[0, const[0]])))
    assert len(face_planes) == 6
    # the cube map's uv layout: every triangle's uvs lie in one cell of the 4 x 3 grid
    uv = flat.tri_uv.numpy()[-12:].reshape(12, 3, 2)
    for t in uv:
        cu = np.floor(t[:, 0].mean() * 4)
        cv = np.floor(t[:, 1].mean() * 3)
        assert ((t[:, 0] >= cu / 4 - 1e-6) & (t[:, 0] <= (cu + 1) / 4 + 1e-6)).all()
        assert ((t[:, 1] >= cv / 3 - 1e-6) & (t[:, 1] <= (cv + 1) / 3 + 1e-6)).all()


@pytest.mark.parametrize("kind", ["sh9", "map"])
def test_uniform_environment_on_convex_object(oracle, kind):
    """Direct light (max_bounce 1) from radiance c in every direction on a convex
    Lambertian object of albedo rho: outgoing radiance rho * c, whatever the
    view (the object hides exactly the lower hemisphere of each face)."""
    c = 1.7
    Vs, Fs = W.icosphere(2)
    sphere = dr.Mesh(torch.tensor(Vs), None, torch.tensor(Fs), None, dr.DiffuseMaterial(torch.tensor([0.5] * 3)))
    if kind == "sh9":
        coeff = torch.zeros((9, 3))
        coeff[0] = c / math.sqrt(1.0 / (4.0 * math.pi))
        env = dr.SH9Light(coeff)
    else:
        env = dr.EnvLight(dr.Texture(torch.full((6, 8, 3), c)))
    cam = H.camera(CPU, 16)
    flat = flatten([sphere], [], CPU, env=env, cameras=[cam])
    integ = dr.PathIntegrator(spp_interior=64, max_bounce=1)
    _, lanes = oracle.render(oracle.Scene(flat), integ.cfg, cam.params(), lane_L=True)
    # lanes whose camera ray hits the sphere (the cube behind it is a light:
    # radiance 0 without enable_light_visable)
    o, d, _ = oracle.camera_rays(cam.params(), 64)
    tri, _, _, _ = oracle.closest_hit(H.tri9(flat), o, d)
    on = (tri >= 0) & (tri < len(Fs))
    assert on.sum() > 2000
    mean = lanes[on].astype(np.float64).mean(0)
    np.testing.assert_allclose(mean, 0.5 * c, rtol=0.02)
