"""GPU parity of the forward path against the CPU oracle (run with -m gpu).

Bar: PCG streams, camera rays, triangle records, edges, sampler tables, LBVH
topology and closest-hit (tri, t, u, v) bit-exact; per-lane and per-pixel
radiance bit-exact with a tolerance fallback (mean rel <= 1e-4, max <= 1e-3).
"""
import ctypes

import numpy as np
import pytest
import torch

import torch_psdr_amd as dr
import torch_psdr_amd._native as N
from torch_psdr_amd import ops
from torch_psdr_amd import workloads as W
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _scene(dev, **kw):
    meshes, lights, leaves = H.objects(dev, **kw)
    res = kw.pop("res", 32) if "res" in kw else 32
    return dr.Scene([H.camera(dev, 32)], meshes, lights)


def _stream():
    return N.current_stream()


def test_pcg_lane_draws(gpu_device, oracle):
    lib = N.load()
    out = torch.empty((4096, 8), dtype=torch.int32, device=gpu_device)
    N.check(lib.psdr_pcg_lane_draws(4190000, 4096, 8, out.data_ptr(), _stream()))
    torch.cuda.synchronize()
    want = oracle.lane_draws(4190000, 4096, 8)
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), want)


@pytest.mark.parametrize("res,spp,kind", [(32, 1, "persp"), (64, 4, "persp"), (32, 2, "ortho")])
def test_camera_rays(gpu_device, oracle, res, spp, kind):
    lib = N.load()
    cp = H.cam_params(res) if kind == "persp" else H.ortho_params(res)
    N_ = spp * res * res
    o = torch.empty((N_, 3), device=gpu_device)
    d = torch.empty((N_, 3), device=gpu_device)
    pix = torch.empty((N_,), dtype=torch.int32, device=gpu_device)
    cam = ops.make_camera(cp)
    N.check(lib.psdr_camera_rays(ctypes.byref(cam), spp, o.data_ptr(), d.data_ptr(), pix.data_ptr(), _stream()))
    torch.cuda.synchronize()
    oo, dd, pp = oracle.camera_rays(cp, spp)
    np.testing.assert_array_equal(o.cpu().numpy(), oo)
    np.testing.assert_array_equal(d.cpu().numpy(), dd)
    np.testing.assert_array_equal(pix.cpu().numpy(), pp)


def _native(scene):
    ns, flat = ops.entry(scene.token)
    return ns, flat


def test_scene_derived_and_samplers(gpu_device, oracle):
    scene = _scene(gpu_device, level=3, tex_res=16)
    ns, flat = _native(scene)
    nt, ne, depth = ns.info()
    lib = ns.lib
    tri7 = torch.empty((nt, 7), device=gpu_device)
    topo = torch.empty((ne, 6), dtype=torch.int32, device=gpu_device)
    geom = torch.empty((ne, 16), device=gpu_device)
    N.check(lib.psdr_scene_derived(ns.handle, tri7.data_ptr(), topo.data_ptr(), geom.data_ptr(), _stream()))
    osc = oracle.Scene(flat)
    t_o, topo_o, geom_o = oracle.prepare(osc)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(tri7.cpu().numpy(), t_o)
    np.testing.assert_array_equal(topo.cpu().numpy(), topo_o)
    np.testing.assert_array_equal(geom.cpu().numpy(), geom_o)
    no = len(flat.objects)
    bufs = [torch.empty(n, device=gpu_device) for n in (no, no, ne, ne, 1, 1)]
    N.check(lib.psdr_scene_samplers(ns.handle, *[b.data_ptr() for b in bufs], _stream()))
    torch.cuda.synchronize()
    want = oracle.samplers(osc, ne)
    for g, w in zip(bufs, want):
        np.testing.assert_array_equal(g.cpu().numpy(), w)


@pytest.mark.parametrize("level", [1, 3, 5])
def test_lbvh_bit_exact(gpu_device, oracle, level):
    scene = _scene(gpu_device, level=level)
    ns, flat = _native(scene)
    n = flat.n_tris
    lib = ns.lib
    morton = torch.empty(n, dtype=torch.int32, device=gpu_device)
    order = torch.empty(n, dtype=torch.int32, device=gpu_device)
    child = torch.empty((n - 1, 2), dtype=torch.int32, device=gpu_device)
    parent = torch.empty(2 * n - 1, dtype=torch.int32, device=gpu_device)
    box = torch.empty((n - 1, 12), device=gpu_device)
    N.check(lib.psdr_scene_bvh(ns.handle, morton.data_ptr(), order.data_ptr(), child.data_ptr(), parent.data_ptr(),
                               box.data_ptr(), _stream()))
    torch.cuda.synchronize()
    m_o, o_o, c_o, p_o, b_o = oracle.lbvh(H.tri9(flat))
    np.testing.assert_array_equal(morton.cpu().numpy().view(np.uint32), m_o)
    np.testing.assert_array_equal(order.cpu().numpy(), o_o)
    np.testing.assert_array_equal(child.cpu().numpy(), c_o)
    np.testing.assert_array_equal(parent.cpu().numpy(), p_o)
    np.testing.assert_array_equal(box.cpu().numpy(), b_o)


@pytest.mark.parametrize("level", [3, 5])
def test_closest_hit_matches_brute_force(gpu_device, oracle, level):
    scene = _scene(gpu_device, level=level)
    ns, flat = _native(scene)
    rng = np.random.default_rng(level)
    n = 200000
    o = rng.uniform(-6, 6, (n, 3)).astype(np.float32)
    tgt = (rng.normal(size=(n, 3)) * np.where(rng.random((n, 1)) < 0.5, 0.6, 4.0)).astype(np.float32)
    d = tgt - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    o[:1000] = [0, 1, 5]  # camera-like origins
    # axis-parallel rays: one or two direction components exactly +-0
    k = np.arange(1000, 31000)
    d[k, k % 3] = np.where(k % 2 == 0, 0.0, -0.0)
    d[k[::3], (k[::3] + 1) % 3] = 0.0
    d[k] /= np.linalg.norm(d[k], axis=1, keepdims=True)
    # rays aimed at mesh vertices from the camera-like origin: grazing hits
    # (barycentrics at a corner), where another triangle test could claim the hit
    Vs = flat.vertices.detach().cpu().numpy()
    kv = np.arange(31000, 41000)
    d[kv] = Vs[rng.integers(0, Vs.shape[0] - 8, kv.size)] - o[kv]
    d[kv] /= np.linalg.norm(d[kv], axis=1, keepdims=True)
    ro, rd = torch.tensor(o, device=gpu_device), torch.tensor(d, device=gpu_device)
    tri = torch.empty(n, dtype=torch.int32, device=gpu_device)
    t, u, v = (torch.empty(n, device=gpu_device) for _ in range(3))
    lib = ns.lib
    N.check(lib.psdr_trace_counters(1, None, 0))
    N.check(lib.psdr_scene_trace(ns.handle, ro.data_ptr(), rd.data_ptr(), n, tri.data_ptr(), t.data_ptr(),
                                 u.data_ptr(), v.data_ptr(), _stream()))
    cnt = (ctypes.c_int64 * N.PSDR_TRACE_COUNTERS)()
    N.check(lib.psdr_trace_counters(0, cnt, N.PSDR_TRACE_COUNTERS))
    torch.cuda.synchronize()
    w_tri, w_t, w_u, w_v = oracle.closest_hit(H.tri9(flat), o, d, brute=(level < 5))
    cls = oracle.hit_classes(H.tri9(flat), o, d)
    g_tri = tri.cpu().numpy()
    mism = g_tri != w_tri
    ties, grazing = int((cls & 1).sum()), int(((cls & 2) > 0).sum())
    print(f"level {level}: {n} rays, hits {(w_tri >= 0).sum()}, id mismatches {mism.sum()}, near ties {ties}, "
          f"grazing {grazing}; GPU: nodes/ray {cnt[1] / cnt[0]:.2f}, triangles/ray {cnt[2] / cnt[0]:.2f}, "
          f"exact tests/ray {cnt[3] / cnt[0]:.2f}")
    # the counting traversal classifies the same rays as the oracle
    assert (cnt[0], cnt[7], cnt[6]) == (n, n, int((w_tri >= 0).sum()))
    assert (cnt[4], cnt[5]) == (ties, grazing)
    assert grazing >= 1000 and ties >= 500  # vertex-aimed rays that reach their vertex (shared by 5-6 triangles)
    # an id may differ from the oracle's only on a near tie or a grazing hit
    # (the contract); t, u, v of every other ray are bit-identical
    assert (cls[mism] != 0).all()
    ok = ~mism
    np.testing.assert_array_equal(t.cpu().numpy()[ok], w_t[ok])
    np.testing.assert_array_equal(u.cpu().numpy()[ok], w_u[ok])
    np.testing.assert_array_equal(v.cpu().numpy()[ok], w_v[ok])


def _compare_images(g, w, tag):
    g = g.astype(np.float64)
    w = w.astype(np.float64)
    exact = (g == w).mean()
    den = np.maximum(np.abs(w), 1e-3)
    rel = np.abs(g - w) / den
    print(f"{tag}: bit-exact {exact:.6f}, mean rel {rel.mean():.3e}, max rel {rel.max():.3e}")
    assert rel.mean() <= 1e-4 and rel.max() <= 1e-3
    return exact


CASES = [
    dict(name="c0-mini", kw=dict(level=3), res=32, cfg=dict(spp_interior=4, max_bounce=2)),
    dict(name="tex-d3", kw=dict(level=2, tex_res=16), res=32, cfg=dict(spp_interior=2, max_bounce=3)),
    dict(name="mis2x2", kw=dict(level=2), res=24, cfg=dict(spp_interior=2, max_bounce=3, mis_light_samples=2,
                                                          mis_bsdf_samples=2)),
    dict(name="nee-only", kw=dict(level=2), res=24, cfg=dict(spp_interior=2, max_bounce=2, mis_bsdf_samples=0)),
    dict(name="bsdf-only", kw=dict(level=2), res=24, cfg=dict(spp_interior=2, max_bounce=2, mis_light_samples=0)),
    dict(name="lights-visible", kw=dict(level=2), res=24, cfg=dict(spp_interior=2, max_bounce=2,
                                                                  enable_light_visable=True)),
    dict(name="bsdf-rough", kw=dict(level=2, bsdf_rough=0.5), res=24, cfg=dict(spp_interior=2, max_bounce=3)),
    dict(name="bsdf-normal-map", kw=dict(level=2, bsdf_rough=0.5, tex_res=8, normal_res=8), res=24,
         cfg=dict(spp_interior=2, max_bounce=3)),
    dict(name="npass3", kw=dict(level=1), res=16, cfg=dict(spp_interior=3, max_bounce=2, n_pass=3)),
    dict(name="ortho-rough", kw=dict(level=2, bsdf_rough=0.5), res=24, cfg=dict(spp_interior=2, max_bounce=3),
         ortho=True),
    dict(name="env-map", kw=dict(level=2, env="map"), res=24, cfg=dict(spp_interior=2, max_bounce=3)),
    dict(name="env-sh9-visible", kw=dict(level=2, env="sh9", bsdf_rough=0.5), res=24,
         cfg=dict(spp_interior=2, max_bounce=3, enable_light_visable=True)),
    dict(name="env-sh9-only", kw=dict(level=2, env="sh9", area_light=False), res=24,
         cfg=dict(spp_interior=2, max_bounce=2)),
    # the reference examples' integrator: 3 light and 3 BSDF samples per depth, roughness 0.5
    dict(name="examples-mis3x3", kw=dict(level=2, bsdf_rough=0.5), res=24,
         cfg=dict(spp_interior=2, max_bounce=3, mis_light_samples=3, mis_bsdf_samples=3)),
    # deeper paths than the configurations use (the depth-8 and general adjoint instances)
    dict(name="deep-d6", kw=dict(level=2), res=16, cfg=dict(spp_interior=2, max_bounce=6)),
    dict(name="deep-d10", kw=dict(level=1), res=16, cfg=dict(spp_interior=2, max_bounce=10)),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_forward_radiance_matches_oracle(gpu_device, oracle, case):
    meshes, lights, _ = H.objects(gpu_device, **case["kw"])
    cam = (H.ortho_camera if case.get("ortho") else H.camera)(gpu_device, case["res"])
    scene = dr.Scene([cam], meshes, lights)
    integ = dr.PathIntegrator(**case["cfg"])
    img = integ.renderC(scene)[0]
    ns, flat = _native(scene)
    N_ = integ.cfg[1] * case["res"] ** 2
    lane = torch.empty((N_, 3), device=gpu_device)
    N.check(ns.lib.psdr_last_lane_radiance(ns.handle, lane.data_ptr(), _stream()))
    torch.cuda.synchronize()
    w_img, w_lane = oracle.render(oracle.Scene(flat), integ.cfg, scene.cameras[0].params(), mode=0, lane_L=True)
    _compare_images(lane.cpu().numpy(), w_lane, case["name"] + " lanes")
    _compare_images(img.cpu().numpy(), w_img, case["name"] + " image")
    assert np.isfinite(img.cpu().numpy()).all() and img.abs().sum() > 0


def test_c0_full_resolution(gpu_device, oracle):
    meshes, lights, _ = H.objects(gpu_device, level=3)
    scene = dr.Scene([H.camera(gpu_device, 256)], meshes, lights)
    integ = dr.PathIntegrator(spp_interior=4, max_bounce=2)
    img = integ.renderC(scene)[0]
    torch.cuda.synchronize()
    ns, flat = _native(scene)
    w = oracle.render(oracle.Scene(flat), integ.cfg, scene.cameras[0].params(), mode=0)
    exact = _compare_images(img.cpu().numpy(), w, "C0 256x256 4spp")
    assert exact > 0.999


@pytest.mark.parametrize("bad", ["past-the-end", "negative", "other-object"])
def test_scene_create_rejects_a_bad_tri_obj(gpu_device, bad):
    """tri_obj entries index the object table in every kernel: an entry outside
    [0, n_objects), or naming an object whose triangle range does not hold the
    triangle, is refused at creation with PSDR_ERR_ARG."""
    from torch_psdr_amd.flat import flatten
    meshes, lights, _ = H.objects(gpu_device, level=1)
    flat = flatten(meshes, lights, gpu_device)
    n_obj = len(meshes) + len(lights)
    t = flat.tri_obj.clone()
    t[3] = {"past-the-end": n_obj, "negative": -1, "other-object": 1}[bad]  # triangle 3 is the sphere's
    flat.tri_obj = t
    keep = []
    desc = flat.desc(keep)
    handle = N.C.c_void_p()
    lib = N.load()
    assert lib.psdr_scene_create(N.C.byref(desc), N.C.byref(handle), _stream()) == N.PSDR_ERR_ARG
    assert b"tri_obj" in lib.psdr_last_error()
    assert not handle
    # the untouched table is accepted
    flat.tri_obj = t.clone().fill_(0)
    flat.tri_obj[:] = flatten(meshes, lights, gpu_device).tri_obj
    desc = flat.desc(keep)
    N.check(lib.psdr_scene_create(N.C.byref(desc), N.C.byref(handle), _stream()))
    lib.psdr_scene_destroy(handle)



@pytest.mark.parametrize("tile", [0, 65536])
def test_cameras_in_one_wavefront_match_single_renders(gpu_device, oracle, tile):
    """renderC over four cameras of different sizes and kinds renders them in
    one wavefront (psdr_render_forward_multi); each image equals the camera's
    own render bit for bit, also with lane tiles that straddle cameras."""
    from torch_psdr_amd import tune
    prev = tune.get_tuning()
    try:
        tune.set_tuning(tile_lanes=tile)
        meshes, lights, _ = H.objects(gpu_device, level=3, tex_res=16, bsdf_rough=0.5)
        cams = [H.camera(gpu_device, 64), H.camera(gpu_device, 96, look_from=(1.5, 1.2, 4.5)),
                H.ortho_camera(gpu_device, 48, look_from=(0.3, 2.5, 4.4)),
                H.camera(gpu_device, 80, look_from=(-2.0, 0.8, 4.6))]
        integ = dr.PathIntegrator(spp_interior=8, max_bounce=3)
        together = integ.renderC(dr.Scene(cams, meshes, lights))
        torch.cuda.synchronize()
        for i, cam in enumerate(cams):
            alone = integ.renderC(dr.Scene([cam], meshes, lights))[0]
            torch.cuda.synchronize()
            assert together[i].shape == alone.shape
            np.testing.assert_array_equal(together[i].cpu().numpy(), alone.cpu().numpy())
            assert float(alone.mean()) > 0
        # one camera against the oracle as well
        scene = dr.Scene([cams[2]], meshes, lights)
        ns, flat = _native(scene)
        ref = oracle.render(oracle.Scene(flat), integ.cfg, cams[2].params(), mode=0)
        _compare_images(together[2].cpu().numpy(), ref, "ortho camera in a 4-camera wavefront")
    finally:
        tune.set_tuning(prev)


def test_forward_with_more_lights_than_the_shading_lds_tables_hold(gpu_device, oracle):
    """Ten area lights (20 emitting triangles): more lights than k_shade stages
    in LDS and more emitting triangles than BSDF-sampled rays test before
    traversing, so the shading kernel reads the scene tables from global
    memory and every BSDF-sampled ray traverses; lanes and image equal the
    oracle's as in the other cases."""
    meshes, lights, _ = H.objects(gpu_device, level=2)
    for k in range(9):
        Vl, Fl = W.quad(0.3, 3.0 + 0.1 * k, facing_down=True)
        Vl = Vl.copy()
        Vl[:, 0] += -2.0 + 0.5 * k
        Vl[:, 2] += 0.25 * (k % 3)
        lights.append(dr.AreaLight(torch.tensor(Vl, device=gpu_device), torch.tensor(Fl, device=gpu_device),
                                   torch.tensor([2.0 + k, 1.0 + 0.5 * k, 3.0], device=gpu_device)))
    cam = H.camera(gpu_device, 24)
    scene = dr.Scene([cam], meshes, lights)
    integ = dr.PathIntegrator(spp_interior=2, max_bounce=3)
    img = integ.renderC(scene)[0]
    ns, flat = _native(scene)
    N_ = integ.cfg[1] * 24 ** 2
    lane = torch.empty((N_, 3), device=gpu_device)
    N.check(ns.lib.psdr_last_lane_radiance(ns.handle, lane.data_ptr(), _stream()))
    torch.cuda.synchronize()
    w_img, w_lane = oracle.render(oracle.Scene(flat), integ.cfg, scene.cameras[0].params(), mode=0, lane_L=True)
    _compare_images(lane.cpu().numpy(), w_lane, "ten lights lanes")
    _compare_images(img.cpu().numpy(), w_img, "ten lights image")
    assert np.isfinite(img.cpu().numpy()).all() and img.abs().sum() > 0
