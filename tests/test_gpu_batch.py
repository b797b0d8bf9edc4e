"""Batched scenes of uneven size (the 64-scene configuration), run with -m gpu.

Scenes share one set of per-render buffers (psdr_scene_share_workspace).  Each
scene's image must equal, bit for bit, the image of the same scene rendered
with buffers of its own, and its gradients the gradients there (to the
reordering of float atomics), whatever the other scenes rendered in between and
whether its forward state was evicted (PSDR_ERR_STATE, then rendered again).
The captured multi-scene step replays to the eager step; at full size (64
scenes, 980-184,320 triangles, 128^2-1024^2, 4-64 spp) the step runs and sampled
lanes of its largest and smallest scenes equal the oracle's.
"""
import ctypes

import numpy as np
import pytest
import torch

import torch_psdr_amd as dr
import torch_psdr_amd._native as N
from torch_psdr_amd import ops
from torch_psdr_amd import workloads as W
from torch_psdr_amd.graph import CapturedStep

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
SMALL = dict(n_scenes=6, seed=4, res_choices=(32, 48), spp_choices=(1, 4), log_tris=(2.5, 3.6), tex_res=16)


def _call_backward(ns, flat, cam, cfg, g):
    gv = torch.zeros_like(flat.vertices)
    gt = torch.zeros_like(flat.texels)
    ge = torch.zeros((max(flat.light_emit.shape[0], 1), 3), device=DEV)
    c, k = ops.make_config(cfg), ops.make_camera(cam)
    rc = ns.lib.psdr_render_backward(ns.handle, ctypes.byref(c), ctypes.byref(k), g.data_ptr(), gv.data_ptr(),
                                     gt.data_ptr(), ge.data_ptr(), None, N.PSDR_TERM_ALL, N.current_stream())
    return rc, gv, gt


def _close(got, want, what):
    tol = 1e-5 * float(want.abs().max()) + 1e-12
    err = float((got - want).abs().max())
    assert err <= tol, (what, err, tol)


def test_shared_buffers_match_dedicated_scenes():
    parts = W.c3_parts(dr, DEV, targets=False, **SMALL)
    scenes = [b.scene(b.vertices, b.tex) for b in parts]
    assert len({id(ops.entry(s.token)[0]) for s in scenes}) == len(parts)
    grads = [torch.randn((b.res, b.res, 3), generator=torch.Generator().manual_seed(30 + i)).to(DEV)
             for i, b in enumerate(parts)]
    # reference: every scene on buffers of its own
    ref = []
    for sc, b, g in zip(scenes, parts, grads):
        _, flat = ops.entry(sc.token)
        own = ops.NativeScene(flat)
        img = ops.forward_native(own, flat, sc.cameras[0].params(), b.integrator.cfg, True)
        rc, gv, gt = _call_backward(own, flat, sc.cameras[0].params(), b.integrator.cfg, g)
        N.check(rc)
        torch.cuda.synchronize()
        ref.append((img, gv, gt))
        del own
    # shared buffers, one scene after another
    for rep in range(2):
        for i, (sc, b, g) in enumerate(zip(scenes, parts, grads)):
            ns, flat = ops.entry(sc.token)
            img = ops.forward_native(ns, flat, sc.cameras[0].params(), b.integrator.cfg, True)
            rc, gv, gt = _call_backward(ns, flat, sc.cameras[0].params(), b.integrator.cfg, g)
            N.check(rc)
            torch.cuda.synchronize()
            print(f"rep {rep} scene {i}: {b.res}^2 x {b.spp} spp, {b.n_tris} tris", flush=True)
            assert torch.equal(img, ref[i][0]), (rep, i)
            _close(gv, ref[i][1], ("vertices", rep, i))
            _close(gt, ref[i][2], ("texels", rep, i))
    # all forwards first: the workspace keeps four forward states, so some
    # backward passes find theirs evicted and ask for the forward again
    imgs = []
    for sc, b in zip(scenes, parts):
        ns, flat = ops.entry(sc.token)
        imgs.append(ops.forward_native(ns, flat, sc.cameras[0].params(), b.integrator.cfg, True))
    evicted = 0
    for i, (sc, b, g) in enumerate(zip(scenes, parts, grads)):
        ns, flat = ops.entry(sc.token)
        rc, gv, gt = _call_backward(ns, flat, sc.cameras[0].params(), b.integrator.cfg, g)
        if rc == N.PSDR_ERR_STATE:
            evicted += 1
            ops.forward_native(ns, flat, sc.cameras[0].params(), b.integrator.cfg, True)
            rc, gv, gt = _call_backward(ns, flat, sc.cameras[0].params(), b.integrator.cfg, g)
        N.check(rc)
        torch.cuda.synchronize()
        assert torch.equal(imgs[i], ref[i][0]), i
        _close(gv, ref[i][1], ("vertices, all forwards first", i))
        _close(gt, ref[i][2], ("texels, all forwards first", i))
    assert evicted >= len(parts) - 4


def test_captured_batched_step_matches_eager():
    parts = W.c3_parts(dr, DEV, **SMALL)

    def step():
        out = []
        for b in parts:
            b.vertices.grad = None
            b.tex.grad = None
            img = b.integrator.renderD(b.scene(b.vertices, b.tex))[0]
            loss = ((img - b.target) ** 2).sum()
            loss.backward()
            out.append((img, b.vertices.grad, b.tex.grad))
        return out

    eager = [tuple(t.detach().clone() for t in r) for r in step()]
    torch.cuda.synchronize()
    cs = CapturedStep(step, warmup=1)
    for rep in range(3):
        out = cs.replay()
        torch.cuda.synchronize()
        print(f"replay {rep}", flush=True)
        for i, ((img, gv, gt), (ei, egv, egt)) in enumerate(zip(out, eager)):
            assert torch.equal(img, ei), (rep, i)
            _close(gv, egv, ("vertices", rep, i))
            _close(gt, egt, ("texels", rep, i))
        # eager renders of other scenes between replays
        with torch.no_grad():
            parts[rep].integrator.renderC(parts[rep].scene(parts[rep].vertices.detach(), parts[rep].tex.detach()))


def test_zero_pdf_extension_keeps_gradients_finite(oracle):
    """Scene 1 of the batched configuration (256^2, 16 spp, 64,980-triangle
    sphere of radius ~2) has a lane whose extension sample has pdf 0: its
    throughput is 0/0 but nothing after it reaches L, so its adjoint is 0.  The
    interior VJP must be finite and equal the oracle's JVP along a random
    vertex + texel tangent."""
    b = W.c3_parts(dr, DEV, n_scenes=2, share=False, targets=False)[1]
    sc = b.scene(b.vertices, b.tex)
    ns, flat = ops.entry(sc.token)
    cam = sc.cameras[0].params()
    ops.forward_native(ns, flat, cam, b.integrator.cfg, True)
    g = torch.randn((b.res, b.res, 3), generator=torch.Generator().manual_seed(41)).to(DEV)
    gv = torch.zeros_like(flat.vertices)
    gt = torch.zeros_like(flat.texels)
    ge = torch.zeros((flat.light_emit.shape[0], 3), device=DEV)
    c, k = ops.make_config(b.integrator.cfg), ops.make_camera(cam)
    N.check(ns.lib.psdr_render_backward(ns.handle, ctypes.byref(c), ctypes.byref(k), g.data_ptr(), gv.data_ptr(),
                                        gt.data_ptr(), ge.data_ptr(), None, N.PSDR_TERM_INTERIOR,
                                        N.current_stream()))
    torch.cuda.synchronize()
    assert torch.isfinite(gv).all() and torch.isfinite(gt).all()
    rng = np.random.default_rng(42)
    wv = np.zeros(flat.vertices.shape, np.float32)
    nv = b.vertices.shape[0]
    wv[:nv] = rng.normal(size=(nv, 3))
    ob = flat.objects[0]
    wt = np.zeros(flat.texels.numel(), np.float32)
    n_tex = 3 * ob["albedo_h"] * ob["albedo_w"]
    wt[ob["albedo_offset"]:ob["albedo_offset"] + n_tex] = rng.normal(size=n_tex)
    _, d_int, _, _ = oracle.render(oracle.Scene(flat, d_vertices=wv.reshape(-1), d_texels=wt), b.integrator.cfg, cam,
                                   mode=1, boundary=False)
    lhs = float((g.cpu().numpy().astype(np.float64) * d_int).sum())
    rhs = float((gv.cpu().numpy().astype(np.float64) * wv).sum() + (gt.cpu().numpy().astype(np.float64) * wt).sum())
    print(f"interior: JVP {lhs:.6e} VJP {rhs:.6e}", flush=True)
    assert np.isfinite(lhs) and abs(lhs - rhs) <= 2e-3 * max(abs(lhs), abs(rhs))


def test_c3_batched_scenes_full_scale(oracle):
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    parts = W.c3_parts(dr, DEV)
    lanes_total = sum(b.res * b.res * b.spp for b in parts)
    print(f"C3: {len(parts)} scenes, {lanes_total / 1e6:.1f} M lanes, tris "
          f"{min(b.n_tris for b in parts)}-{max(b.n_tris for b in parts)}", flush=True)
    for i, b in enumerate(parts):
        img = b.integrator.renderD(b.scene(b.vertices, b.tex))[0]
        ((img - b.target) ** 2).mean().backward()
        torch.cuda.synchronize()
        assert torch.isfinite(img).all() and img.mean() > 0, i
        for name, t in (("vertices", b.vertices), ("texels", b.tex)):
            assert t.grad is not None and torch.isfinite(t.grad).all() and t.grad.abs().sum() > 0, (i, name)
    used = free0 - torch.cuda.mem_get_info()[0]
    print(f"C3: device memory in use after the step {used / 2**30:.1f} GiB", flush=True)
    # per-lane radiance of the largest and the smallest scene vs the oracle
    order = sorted(range(len(parts)), key=lambda i: parts[i].res ** 2 * parts[i].spp * parts[i].n_tris)
    for i in (order[-1], order[0]):
        b = parts[i]
        sc = b.scene(b.vertices.detach(), b.tex.detach())
        ns, flat = ops.entry(sc.token)
        cam = sc.cameras[0].params()
        ops.forward_native(ns, flat, cam, b.integrator.cfg, False)
        n_lanes = b.res * b.res * b.spp
        lanes = torch.empty((n_lanes, 3), device=DEV)
        N.check(ns.lib.psdr_last_lane_radiance(ns.handle, lanes.data_ptr(), N.current_stream()))
        torch.cuda.synchronize()
        # 32 pixels at the centre row (through the sphere) and 32 at a quarter height
        rows = (b.res // 2, b.res // 4)
        idx = np.concatenate([np.arange((r * b.res + b.res // 2 - 16) * b.spp, (r * b.res + b.res // 2 + 16) * b.spp)
                              for r in rows])
        want = oracle.render_lanes(oracle.Scene(flat), b.integrator.cfg, cam, idx).astype(np.float64)
        got = lanes[torch.tensor(idx, device=DEV)].cpu().numpy().astype(np.float64)
        rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-3)
        print(f"C3 scene {i} ({b.res}^2 x {b.spp} spp, {b.n_tris} tris): bit-exact {(got == want).mean():.4f}, "
              f"mean rel {rel.mean():.2e}, max rel {rel.max():.2e}", flush=True)
        assert want.sum() > 0
        assert rel.mean() <= 1e-4 and rel.max() <= 1e-3
