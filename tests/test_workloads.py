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
