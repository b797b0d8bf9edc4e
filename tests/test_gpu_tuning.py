"""Launch shapes (psdr_set_tuning) and the autotuner leave results unchanged (run with -m gpu).

Images must be bit-identical under every shape (each pixel sums its own lanes
in lane order), and gradients equal up to the order of their float atomics.
"""
import numpy as np
import pytest
import torch

import torch_psdr_amd as dr
from torch_psdr_amd import tune
from tests import helpers as H

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(trace_blocks_per_cu=1, trace_stack_lds=8, shade_blocks_per_cu=1, adjoint_blocks_per_cu=1,
         direct_blocks_per_cu=1),
    dict(trace_blocks_per_cu=64, trace_stack_lds=32, shade_blocks_per_cu=64, adjoint_blocks_per_cu=64,
         direct_blocks_per_cu=64),
    dict(trace_blocks_per_cu=4, trace_stack_lds=16, shade_blocks_per_cu=2, adjoint_blocks_per_cu=8,
         direct_blocks_per_cu=4),
]


@pytest.fixture
def restore_tuning():
    prev = tune.get_tuning()
    yield
    tune.set_tuning(prev)


def _step(dev, res=48):
    meshes, lights, leaves = H.objects(dev, level=3, tex_res=16, requires_grad=True)
    scene = dr.Scene([H.camera(dev, res)], meshes, lights)
    integ = dr.PathIntegrator(spp_interior=4, max_bounce=3, spp_primary_edge=2, spp_secondary_edge=2)
    img = integ.renderD(scene)[0]
    loss = (img * torch.linspace(0.5, 1.5, img.numel(), device=dev).view_as(img)).sum()
    loss.backward()
    torch.cuda.synchronize()
    return img.detach().cpu().numpy(), leaves["sphere_v"].grad.cpu().numpy(), leaves["color"].grad.cpu().numpy()


def test_shapes_do_not_change_results(gpu_device, restore_tuning):
    tune.set_tuning(tune.DEFAULTS)
    img0, gv0, gt0 = _step(gpu_device)
    assert np.abs(gv0).max() > 0 and np.abs(gt0).max() > 0
    for shape in SHAPES:
        print("shape", shape, flush=True)
        tune.set_tuning(shape)
        img, gv, gt = _step(gpu_device)
        np.testing.assert_array_equal(img, img0)
        np.testing.assert_allclose(gv, gv0, rtol=1e-4, atol=1e-5 * np.abs(gv0).max())
        np.testing.assert_allclose(gt, gt0, rtol=1e-4, atol=1e-5 * np.abs(gt0).max())


def test_autotune_picks_from_the_space_and_applies_it(gpu_device, restore_tuning):
    tune.set_tuning(tune.DEFAULTS)
    meshes, lights, _ = H.objects(gpu_device, level=3, tex_res=16, requires_grad=True)
    integ = dr.PathIntegrator(spp_interior=4, max_bounce=2, spp_primary_edge=1, spp_secondary_edge=1)
    cam = H.camera(gpu_device, 64)

    def step():
        scene = dr.Scene([cam], meshes, lights)
        img = integ.renderD(scene)[0]
        (img ** 2).sum().backward()

    # each field's space holds its default (the start of the sweep)
    space = {"trace_blocks_per_cu": (8, 16), "trace_stack_lds": (8, 16), "shade_blocks_per_cu": (4, 8)}
    assert all(tune.DEFAULTS[k] in vals for k, vals in space.items())
    res = tune.autotune(step, space=space, repeats=2, rounds=1)
    best = res["tuning"]
    assert tune.get_tuning() == best
    for k, vals in space.items():
        assert best[k] in vals
    assert best["adjoint_blocks_per_cu"] == tune.DEFAULTS["adjoint_blocks_per_cu"]
    # the start plus every value other than the current best of each field
    assert res["ms"] > 0 and len(res["trials"]) >= 1 + 3
    assert res["ms"] == min(ms for _, ms in res["trials"] if _ == best)
    for shape, _ in res["trials"]:
        assert all(shape[k] in vals for k, vals in space.items())
    key = tune.workload_key(dr.Scene([cam], meshes, lights), integ)
    assert "lanes16384" in key and "tris" in key


def _tile_step(dev, res=96, spp=16):
    meshes, lights, leaves = H.objects(dev, level=3, tex_res=16, requires_grad=True)
    scene = dr.Scene([H.camera(dev, res)], meshes, lights)
    integ = dr.PathIntegrator(spp_interior=spp, max_bounce=3, spp_primary_edge=1, spp_secondary_edge=1)
    img = integ.renderD(scene)[0]
    loss = (img * torch.linspace(0.5, 1.5, img.numel(), device=dev).view_as(img)).sum()
    loss.backward()
    torch.cuda.synchronize()
    import ctypes
    import torch_psdr_amd._native as N
    from torch_psdr_amd import ops
    ns, _ = ops.entry(scene.token)
    counts = (ctypes.c_int64 * 5)()
    traced = ctypes.c_int64()
    N.check(ns.lib.psdr_last_render_stats(ns.handle, counts, 5, ctypes.byref(traced)))
    return (img.detach().cpu().numpy(), leaves["sphere_v"].grad.cpu().numpy(), leaves["color"].grad.cpu().numpy(),
            list(counts), traced.value)


def test_lane_tiles_do_not_change_results(gpu_device, restore_tuning):
    """The forward render in lane tiles (147,456 lanes in tiles of 65,536 and
    100,000: the last tile partial) gives the untiled image bit for bit, the
    same per-depth path counts and rays, and the same gradients up to the
    order of their float atomics."""
    tune.set_tuning(tune.DEFAULTS)
    img0, gv0, gt0, c0, r0 = _tile_step(gpu_device)
    assert c0[0] == 96 * 96 * 16 and r0 > c0[0]
    for tile in (65536, 100000):
        print("tile", tile, flush=True)
        tune.set_tuning(tile_lanes=tile)
        img, gv, gt, c, r = _tile_step(gpu_device)
        np.testing.assert_array_equal(img, img0)
        assert (c, r) == (c0, r0)
        np.testing.assert_allclose(gv, gv0, rtol=1e-4, atol=1e-5 * np.abs(gv0).max())
        np.testing.assert_allclose(gt, gt0, rtol=1e-4, atol=1e-5 * np.abs(gt0).max())
