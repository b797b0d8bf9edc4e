"""Configuration scenes (CPU): the terrain builder of the large-scene
configuration and the oracle's lane-range render used to check it at full size."""
import numpy as np
import torch

import torch_psdr_amd as dr
from torch_psdr_amd import workloads as W
from torch_psdr_amd.flat import flatten

CPU = torch.device("cpu")


def test_terrain_grid_and_noise():
    V, F, uv = W.terrain_grid(5, size=4.0)
    assert V.shape == (25, 3) and F.shape == (32, 3) and uv.shape == (25, 2)
    # every cell split into two triangles of half the cell area, facing +y
    P = V[F]
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    np.testing.assert_allclose(np.abs(n[:, 1]) / 2, 0.5, rtol=1e-6)
    assert len({tuple(sorted(f)) for f in F.tolist()}) == 32
    h = W.fbm_value_noise(64)
    h2 = W.fbm_value_noise(64)
    np.testing.assert_array_equal(h, h2)  # seeded
    assert h.dtype == np.float32 and 0.2 < h.std() < 1.0 and np.abs(h).max() < 2.0
    # smooth at the coarsest scale: neighbour differences far below the range
    assert np.abs(np.diff(h, axis=0)).mean() < 0.1 * (h.max() - h.min())


def test_lane_range_render_matches_full_render(oracle):
    p = W.c4_parts(dr, CPU, n=24, res=12, spp=2, depth=3, tex_res=8)
    cams, meshes, lights = p["objects"](p["heights"], p["tex"])
    flat = flatten(meshes, lights, CPU)
    cfg, cp = p["integrator"].cfg, cams[0].params()
    img, lanes = oracle.render(oracle.Scene(flat), cfg, cp, lane_L=True)
    assert img.sum() > 0
    idx = np.r_[37:138, 200:210]
    part = oracle.render_lanes(oracle.Scene(flat), cfg, cp, idx)
    np.testing.assert_array_equal(part, lanes[idx])


def test_geodesic_sphere_is_a_closed_outward_unit_sphere():
    for k in (1, 2, 5, 13):
        V, F = W.geodesic_sphere(k)
        assert F.shape == (20 * k * k, 3) and V.shape == (10 * k * k + 2, 3)
        # every edge shared by exactly two triangles, in opposite directions
        directed = {(int(a), int(b)) for f in F for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0]))}
        assert len(directed) == 3 * F.shape[0]
        assert all((b, a) in directed for a, b in directed)
        P = V[F].astype(np.float64)
        n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
        assert (np.einsum("ij,ij->i", n, P.mean(1)) > 0).all()
        np.testing.assert_allclose(np.linalg.norm(V, axis=1), 1.0, rtol=1e-6)
        # k^2 triangles per icosahedron face: total area tends to 4 pi
        if k == 13:
            assert abs(0.5 * np.linalg.norm(n, axis=1).sum() - 4 * np.pi) < 0.01 * 4 * np.pi
    # frequency 1 is the icosahedron itself
    V1, F1 = W.geodesic_sphere(1)
    V0, F0 = W.icosphere(0)
    np.testing.assert_allclose(np.sort(V1, 0), np.sort(V0, 0), atol=1e-7)


def test_batched_configuration_draws():
    sp = W.c3_specs()
    assert len(sp) == 64
    for a, b in zip(sp, W.c3_specs()):  # seeded
        assert (a["k"], a["res"], a["spp"], a["scale"], a["noise_seed"]) == (b["k"], b["res"], b["spp"], b["scale"],
                                                                             b["noise_seed"])
        np.testing.assert_array_equal(a["rotation"], b["rotation"])
    n = np.array([s["n_tris"] for s in sp])
    assert n.min() >= 980 and n.max() <= 20 * 102 ** 2          # 20 k^2 around 10^[3, 5.3]
    assert n.max() > 10 * n.min()
    assert {s["res"] for s in sp} == {128, 256, 512, 1024}
    assert {s["spp"] for s in sp} == {4, 16, 64}
    for s in sp:
        Q = s["rotation"]
        np.testing.assert_allclose(Q @ Q.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(Q) - 1.0) < 1e-12 and 0.5 <= s["scale"] <= 2.0
