"""GPU adjoint parity: VJP(GPU) . w == g . JVP(oracle) per term (run with -m gpu).

The oracle's forward-mode derivative restates renderD's autodiff semantics
(every detach point of src/path_integrator.cc); the GPU computes the reverse
mode by hand-written adjoint kernels.  For a random image adjoint g and random
parameter directions w the two scalars must agree to fp32 accumulation error.
"""
import ctypes

import numpy as np
import pytest
import torch

import torch_psdr_amd as dr
import torch_psdr_amd._native as N
from torch_psdr_amd import ops
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _setup(dev, kw, res, cfg, ortho=False):
    meshes, lights, leaves = H.objects(dev, requires_grad=True, **kw)
    cam = (H.ortho_camera if ortho else H.camera)(dev, res)
    scene = dr.Scene([cam], meshes, lights)
    integ = dr.PathIntegrator(**cfg)
    return scene, integ


def _gpu_vjp(scene, integ, g, terms):
    ns, flat = ops.entry(scene.token)
    cam = scene.cameras[0].params()
    ops.forward_native(ns, flat, cam, integ.cfg, True)
    gv = torch.zeros_like(flat.vertices)
    gt = torch.zeros_like(flat.texels)
    ge = torch.zeros((max(flat.light_emit.shape[0], 1), 3), device=gv.device)
    gc = torch.zeros(N.PSDR_CAMERA_GRAD, device=gv.device)
    c, k = ops.make_config(integ.cfg), ops.make_camera(cam)
    N.check(ns.lib.psdr_render_backward(ns.handle, ctypes.byref(c), ctypes.byref(k), g.data_ptr(), gv.data_ptr(),
                                        gt.data_ptr(), ge.data_ptr(), gc.data_ptr(), terms, N.current_stream()))
    torch.cuda.synchronize()
    return gv.cpu().numpy().astype(np.float64), gt.cpu().numpy().astype(np.float64), \
        ge.cpu().numpy()[: flat.light_emit.shape[0]].astype(np.float64), gc.cpu().numpy().astype(np.float64)


def _check(oracle, scene, integ, grads, g, term, block, seed, tol):
    ns, flat = ops.entry(scene.token)
    rng = np.random.default_rng(seed)
    wv = rng.normal(size=tuple(flat.vertices.shape)).astype(np.float32) if block in ("v", "all") else None
    wt = rng.normal(size=(flat.texels.numel(),)).astype(np.float32) if block in ("t", "all") else None
    we = rng.normal(size=tuple(flat.light_emit.shape)).astype(np.float32) if block in ("e", "all") else None
    wc = rng.normal(size=N.PSDR_CAMERA_GRAD).astype(np.float32) if block in ("c", "all") else None
    osc = oracle.Scene(flat, d_vertices=wv, d_texels=wt, d_emit=we)
    img, d_int, d_pri, d_dir = oracle.render(osc, integ.cfg, scene.cameras[0].params(), mode=1,
                                             boundary=(term != N.PSDR_TERM_INTERIOR), d_camera=wc)
    d = {N.PSDR_TERM_INTERIOR: d_int, N.PSDR_TERM_PRIMARY: d_pri, N.PSDR_TERM_DIRECT: d_dir}[term]
    lhs = float((g.cpu().numpy().astype(np.float64) * d.astype(np.float64)).sum())
    gv, gt, ge, gc = grads
    rhs = 0.0
    if wv is not None:
        rhs += float((gv * wv).sum())
    if wt is not None:
        rhs += float((gt * wt).sum())
    if we is not None:
        rhs += float((ge * we).sum())
    if wc is not None:
        rhs += float((gc * wc).sum())
    scale = max(abs(lhs), abs(rhs), 1e-12)
    print(f"term {term} block {block}: JVP {lhs:.6e}  VJP {rhs:.6e}  rel {abs(lhs - rhs) / scale:.2e}")
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs)
    return lhs


INTERIOR_CASES = [
    dict(name="tex-d2", kw=dict(level=2, tex_res=8), res=24, cfg=dict(spp_interior=2, max_bounce=2)),
    dict(name="const-d3", kw=dict(level=2), res=24, cfg=dict(spp_interior=2, max_bounce=3)),
    dict(name="mis2x2", kw=dict(level=1), res=16, cfg=dict(spp_interior=2, max_bounce=3, mis_light_samples=2,
                                                          mis_bsdf_samples=2)),
    dict(name="bsdf-rough", kw=dict(level=2, bsdf_rough=0.5, tex_res=8), res=16,
         cfg=dict(spp_interior=2, max_bounce=2)),
    dict(name="elv", kw=dict(level=1), res=16, cfg=dict(spp_interior=2, max_bounce=2, enable_light_visable=True)),
    dict(name="normal-map", kw=dict(level=2, bsdf_rough=0.5, tex_res=8, normal_res=8), res=16,
         cfg=dict(spp_interior=2, max_bounce=2)),
    dict(name="env-map", kw=dict(level=2, env="map"), res=16, cfg=dict(spp_interior=2, max_bounce=2)),
    dict(name="env-sh9", kw=dict(level=2, env="sh9", bsdf_rough=0.5), res=16,
         cfg=dict(spp_interior=2, max_bounce=2, enable_light_visable=True)),
    dict(name="examples-mis3x3", kw=dict(level=1, bsdf_rough=0.5, tex_res=8), res=16,
         cfg=dict(spp_interior=2, max_bounce=3, mis_light_samples=3, mis_bsdf_samples=3)),
    dict(name="deep-d6", kw=dict(level=1, tex_res=8), res=16, cfg=dict(spp_interior=2, max_bounce=6)),
    dict(name="deep-d10", kw=dict(level=1), res=12, cfg=dict(spp_interior=2, max_bounce=10)),
]


@pytest.mark.parametrize("case", INTERIOR_CASES, ids=[c["name"] for c in INTERIOR_CASES])
def test_interior_vjp_matches_oracle_jvp(gpu_device, oracle, case):
    scene, integ = _setup(gpu_device, case["kw"], case["res"], case["cfg"])
    g = torch.randn((case["res"], case["res"], 3), generator=torch.Generator().manual_seed(7)).to(gpu_device)
    grads = _gpu_vjp(scene, integ, g, N.PSDR_TERM_INTERIOR)
    assert np.isfinite(grads[0]).all() and np.isfinite(grads[1]).all()
    for i, block in enumerate(["v", "t", "e"]):
        _check(oracle, scene, integ, grads, g, N.PSDR_TERM_INTERIOR, block, 100 + i, 2e-3)


@pytest.mark.parametrize("term,env,mis", [(N.PSDR_TERM_PRIMARY, None, 1), (N.PSDR_TERM_DIRECT, None, 1),
                                          (N.PSDR_TERM_PRIMARY, "sh9", 1), (N.PSDR_TERM_DIRECT, "map", 1),
                                          (N.PSDR_TERM_PRIMARY, None, 3), (N.PSDR_TERM_DIRECT, None, 3)],
                         ids=["primary", "direct", "primary-env-sh9", "direct-env-map", "primary-mis3",
                              "direct-mis3"])
def test_boundary_vjp_matches_oracle_jvp(gpu_device, oracle, term, env, mis):
    # mis 3: the reference examples' light-sample count (the primary term's
    # preprocessing then runs one cell per thread instead of four lanes per cell)
    scene, integ = _setup(gpu_device, dict(level=2, env=env), 32,
                          dict(spp_interior=1, max_bounce=2, spp_primary_edge=2, spp_secondary_edge=2,
                               mis_light_samples=mis, mis_bsdf_samples=mis))
    g = torch.randn((32, 32, 3), generator=torch.Generator().manual_seed(11)).to(gpu_device)
    grads = _gpu_vjp(scene, integ, g, term)
    assert np.isfinite(grads[0]).all()
    assert np.abs(grads[0]).sum() > 0
    _check(oracle, scene, integ, grads, g, term, "v", 200, 2e-3)


# Camera parameters reach the image through the camera vertex's wi (the
# DiffuseBsdf's roughness term) and the primary boundary's origin term; an
# orthographic origin also moves with the edge point (vertex block).
CAMERA_CASES = [
    dict(name="interior-persp", term=N.PSDR_TERM_INTERIOR, ortho=False),
    dict(name="interior-env-sh9-visible", term=N.PSDR_TERM_INTERIOR, ortho=False, env="sh9"),
    dict(name="interior-ortho", term=N.PSDR_TERM_INTERIOR, ortho=True),
    dict(name="primary-persp", term=N.PSDR_TERM_PRIMARY, ortho=False),
    dict(name="primary-ortho", term=N.PSDR_TERM_PRIMARY, ortho=True),
    dict(name="direct-ortho", term=N.PSDR_TERM_DIRECT, ortho=True),
]


@pytest.mark.parametrize("case", CAMERA_CASES, ids=[c["name"] for c in CAMERA_CASES])
def test_camera_vjp_matches_oracle_jvp(gpu_device, oracle, case):
    term = case["term"]
    env = case.get("env")
    scene, integ = _setup(gpu_device, dict(level=2, bsdf_rough=0.5, env=env), 24,
                          dict(spp_interior=2, max_bounce=2, spp_primary_edge=2, spp_secondary_edge=2,
                               enable_light_visable=env is not None),
                          ortho=case["ortho"])
    g = torch.randn((24, 24, 3), generator=torch.Generator().manual_seed(13)).to(gpu_device)
    grads = _gpu_vjp(scene, integ, g, term)
    assert np.isfinite(grads[3]).all()
    if term != N.PSDR_TERM_DIRECT:
        assert np.abs(grads[3]).sum() > 0
        _check(oracle, scene, integ, grads, g, term, "c", 300, 2e-3)
    else:
        assert np.abs(grads[3]).sum() == 0  # the shadow term does not see the camera
    if term != N.PSDR_TERM_INTERIOR:
        _check(oracle, scene, integ, grads, g, term, "v", 301, 2e-3)


def test_renderD_autograd_end_to_end(gpu_device, oracle):
    meshes, lights, leaves = H.objects(gpu_device, level=2, tex_res=8, requires_grad=True)
    scene = dr.Scene([H.camera(gpu_device, 24)], meshes, lights)
    integ = dr.PathIntegrator(spp_interior=2, max_bounce=2)
    img = integ.renderD(scene)[0]
    target = torch.zeros_like(img)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    for name in ("sphere_v", "color", "emit", "light_v"):
        gr = leaves[name].grad
        assert gr is not None and torch.isfinite(gr).all(), name
    assert leaves["color"].grad.abs().sum() > 0
    assert leaves["sphere_v"].grad.abs().sum() > 0
    # renderC and renderD agree in value
    img_c = integ.renderC(scene)[0]
    np.testing.assert_allclose(img.detach().cpu().numpy(), img_c.cpu().numpy(), rtol=0, atol=0)


def test_all_terms_equal_sum_of_terms(gpu_device):
    """PSDR_TERM_ALL runs the three terms side by side on forked streams; its
    gradients must equal the sum of the terms computed one call each."""
    scene, integ = _setup(gpu_device, dict(level=2, tex_res=8), 32,
                          dict(spp_interior=2, max_bounce=2, spp_primary_edge=2, spp_secondary_edge=2))
    g = torch.randn((32, 32, 3), generator=torch.Generator().manual_seed(5)).to(gpu_device)
    parts = [_gpu_vjp(scene, integ, g, t) for t in (N.PSDR_TERM_INTERIOR, N.PSDR_TERM_PRIMARY, N.PSDR_TERM_DIRECT)]
    for rep in range(2):
        tot = _gpu_vjp(scene, integ, g, N.PSDR_TERM_ALL)
        for b in range(4):
            ref = sum(p[b] for p in parts)
            scale = max(np.abs(ref).max(), 1e-12)
            assert np.abs(tot[b] - ref).max() <= 1e-4 * scale, (rep, b, np.abs(tot[b] - ref).max(), scale)


@pytest.mark.parametrize("n_light", [1, 2, 0])
def test_plain_launch_shape_matches_fused(gpu_device, n_light):
    """The reference-shaped baseline (no compaction, one synchronised trace per
    ray kind, unsorted boundary samples, sequential terms, the primary edges'
    hypercube preprocessing as wavefront launches) renders the same image bit
    for bit and the same gradients up to accumulation order.  The fused step
    runs that preprocessing as one kernel (a lane per ray of a depth with one
    light sample per vertex, a thread per cell otherwise); a different cell
    mass would move the boundary samples and change the gradients."""
    meshes, lights, leaves = H.objects(gpu_device, level=2, tex_res=8, requires_grad=True)
    scene = dr.Scene([H.camera(gpu_device, 24)], meshes, lights)
    integ = dr.PathIntegrator(spp_interior=2, max_bounce=3, spp_primary_edge=2, spp_secondary_edge=2,
                              mis_light_samples=n_light)
    target = torch.rand((24, 24, 3), generator=torch.Generator().manual_seed(3)).to(gpu_device)
    out = {}
    for shape in ("fused", "plain", "fused"):
        for v in leaves.values():
            v.grad = None
        with ops.launch_shape(shape):
            img = integ.renderD(scene)[0]
            ((img - target) ** 2).sum().backward()
        torch.cuda.synchronize()
        out.setdefault(shape, []).append((img.detach().cpu().numpy(),
                                          {k: v.grad.cpu().numpy() for k, v in leaves.items()}))
    img_f, g_f = out["fused"][0]
    img_p, g_p = out["plain"][0]
    np.testing.assert_array_equal(img_f, img_p)
    np.testing.assert_array_equal(img_f, out["fused"][1][0])
    for k in g_f:
        scale = max(np.abs(g_f[k]).max(), 1e-12)
        assert np.abs(g_f[k] - g_p[k]).max() <= 1e-4 * scale, k


def test_scene_update_refreshes_gradient_masks(gpu_device):
    """A native scene is shared by Scenes of the same topology; a Scene whose
    tensors require gradients after one whose tensors did not must still get
    them (the object table, with its gradient masks, is re-uploaded)."""
    meshes0, lights0, _ = H.objects(gpu_device, level=2, tex_res=8, requires_grad=False)
    integ = dr.PathIntegrator(spp_interior=2, max_bounce=2)
    with torch.no_grad():
        integ.renderC(dr.Scene([H.camera(gpu_device, 24)], meshes0, lights0))
    meshes, lights, leaves = H.objects(gpu_device, level=2, tex_res=8, requires_grad=True)
    img = integ.renderD(dr.Scene([H.camera(gpu_device, 24)], meshes, lights))[0]
    (img ** 2).sum().backward()
    assert leaves["color"].grad.abs().sum() > 0
    assert leaves["sphere_v"].grad.abs().sum() > 0
    assert leaves["emit"].grad.abs().sum() > 0


def test_multi_camera_renderD_matches_single_camera_sum(gpu_device):
    """renderD over several cameras (more than the scene keeps forward states
    for) gives each camera's image and the sum of the single-camera gradients,
    camera parameters included; a perspective camera's look_at, up and vfov get
    no gradient (its origin is its position)."""
    meshes, lights, leaves = H.objects(gpu_device, level=2, bsdf_rough=0.5, tex_res=8, requires_grad=True)
    views = [(0.0, 1.0, 5.0), (1.5, 1.2, 4.5), (-2.0, 0.8, 4.6), (0.3, 2.5, 4.4), (0.0, 0.6, 5.2)]
    cams = [H.camera(gpu_device, 16, look_from=v, requires_grad=True) for v in views[:3]] + \
        [H.ortho_camera(gpu_device, 16, look_from=v, requires_grad=True) for v in views[3:]]
    integ = dr.PathIntegrator(spp_interior=2, max_bounce=2, spp_primary_edge=2, spp_secondary_edge=2)
    targets = [torch.rand((16, 16, 3), generator=torch.Generator().manual_seed(20 + i)).to(gpu_device)
               for i in range(len(cams))]
    cam_leaves = [(c.look_from, c.look_at, c.up, c._extent) for c in cams]

    def grads():
        out = {k: v.grad.detach().cpu().numpy().copy() for k, v in leaves.items()}
        for i, cl in enumerate(cam_leaves):
            out[f"cam{i}"] = np.concatenate([t.grad.detach().cpu().numpy().reshape(-1) for t in cl])
        return out

    def clear():
        for v in list(leaves.values()) + [t for cl in cam_leaves for t in cl]:
            v.grad = None

    clear()
    imgs = integ.renderD(dr.Scene(cams, meshes, lights))
    sum(((im - t) ** 2).sum() for im, t in zip(imgs, targets)).backward()
    torch.cuda.synchronize()
    together = grads()
    ref = None
    for i, cam in enumerate(cams):
        clear()
        for cl in cam_leaves:
            for t in cl:
                t.grad = torch.zeros_like(t)
        img = integ.renderD(dr.Scene([cam], meshes, lights))[0]
        np.testing.assert_array_equal(img.detach().cpu().numpy(), imgs[i].detach().cpu().numpy())
        ((img - targets[i]) ** 2).sum().backward()
        torch.cuda.synchronize()
        g = grads()
        ref = g if ref is None else {k: ref[k] + g[k] for k in ref}
    for k in ref:
        scale = max(np.abs(ref[k]).max(), 1e-12)
        assert np.abs(together[k] - ref[k]).max() <= 1e-4 * scale, (k, np.abs(together[k] - ref[k]).max(), scale)
    for i in range(3):
        assert np.abs(together[f"cam{i}"][:3]).sum() > 0
        np.testing.assert_array_equal(together[f"cam{i}"][3:], 0.0)
    for i in range(3, 5):
        assert np.isfinite(together[f"cam{i}"]).all() and np.abs(together[f"cam{i}"]).sum() > 0
