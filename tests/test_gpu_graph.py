"""HIP-graph capture of the optimisation step (torch_psdr_amd.graph.CapturedStep).

A captured step (Scene rebuild with its LBVH, renderD, loss, backward over the
interior and both boundary terms, optimizer update) must give on EVERY replay
what the eager step gives for the same parameters, with eager renders and host
synchronisations between the replays: the image bit for bit (the forward is
deterministic), the gradients to the reordering of their float atomics.
"""
import pytest
import torch

import torch_psdr_amd as dr
from torch_psdr_amd import ops
from torch_psdr_amd import workloads as W
from torch_psdr_amd.graph import CapturedStep

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _problem():
    # the 512^2 vertex + texture step's scene at 64^2, 4 spp, level-3 sphere, 32^2 texture
    P = W.c2_parts(dr, DEV, res=64, spp=4, spp_edge=1, level=3, tex_res=32)
    with torch.no_grad():
        target = P["integrator"].renderC(P["scene"](P["V0"], P["tex_target"]))[0]
    return P, target


def _eager(P, target, verts, tex):
    """Image, loss and gradients of one eager step at the given parameter values."""
    v = verts.detach().clone().requires_grad_(True)
    t = tex.detach().clone().requires_grad_(True)
    img = P["integrator"].renderD(P["scene"](v, t))[0]
    loss = ((img - target) ** 2).sum()
    loss.backward()
    return img.detach(), loss.detach(), v.grad, t.grad


def _close_grad(got, want):
    # float atomics accumulate in a different order on every run
    tol = 1e-5 * float(want.abs().max()) + 1e-12
    err = float((got - want).abs().max())
    assert err <= tol, (err, tol)


def test_every_replay_matches_the_eager_step():
    P, target = _problem()
    verts = P["Vp"].clone().requires_grad_(True)
    tex = P["tex_opt"].clone().requires_grad_(True)

    def step():
        verts.grad = None
        tex.grad = None
        img = P["integrator"].renderD(P["scene"](verts, tex))[0]
        loss = ((img - target) ** 2).sum()
        loss.backward()
        return img, loss, verts.grad, tex.grad

    n_entries = len(ops._entries)
    ref = _eager(P, target, verts, tex)
    # a finished step leaves no scene registered (nor its autograd graph alive)
    assert len(ops._entries) == n_entries
    cs = CapturedStep(step, warmup=2)
    for r in range(4):
        img, loss, gv, gt = cs.replay()
        torch.cuda.synchronize()
        print(f"replay {r}", flush=True)
        assert torch.equal(img, ref[0]), r
        assert torch.equal(loss, ref[1]), (r, float(loss), float(ref[1]))
        _close_grad(gv, ref[2])
        _close_grad(gt, ref[3])
        # eager work on the same native scene between replays: a renderC and a
        # full eager step at the same sizes
        with torch.no_grad():
            again = P["integrator"].renderC(P["scene"](verts.detach(), tex.detach()))[0]
        assert torch.equal(again, ref[0])
        e = _eager(P, target, verts, tex)
        torch.cuda.synchronize()
        assert torch.equal(e[0], ref[0])


def test_captured_optimisation_follows_the_changing_geometry():
    """With an SGD update inside the graph the parameters change on every
    replay; the graph's Scene rebuild (triangle records, edges, samplers, LBVH)
    must follow them: each replay's image equals an eager render at the
    parameters it started from, its gradients the eager ones there."""
    P, target = _problem()
    verts = P["Vp"].clone().requires_grad_(True)
    tex = P["tex_opt"].clone().requires_grad_(True)
    _, _, gv0, gt0 = _eager(P, target, verts, tex)
    # steps that move the vertices by ~1e-3 and the texels by ~1e-2 each
    opt = torch.optim.SGD([{"params": [verts], "lr": 1e-3 / float(gv0.abs().max())},
                           {"params": [tex], "lr": 1e-2 / float(gt0.abs().max())}])

    def step():
        opt.zero_grad(set_to_none=True)
        img = P["integrator"].renderD(P["scene"](verts, tex))[0]
        loss = ((img - target) ** 2).sum()
        loss.backward()
        opt.step()
        return img, verts.grad, tex.grad

    v0, t0 = verts.detach().clone(), tex.detach().clone()
    cs = CapturedStep(step, warmup=1)
    with torch.no_grad():
        verts.copy_(v0)
        tex.copy_(t0)
    for r in range(3):
        vk, tk = verts.detach().clone(), tex.detach().clone()
        img, gv, gt = cs.replay()
        torch.cuda.synchronize()
        print(f"replay {r}", flush=True)
        e_img, _, e_gv, e_gt = _eager(P, target, vk, tk)
        torch.cuda.synchronize()
        assert torch.equal(img, e_img), r
        _close_grad(gv, e_gv)
        _close_grad(gt, e_gt)
        moved = float((verts.detach() - vk).abs().max())
        assert moved > 1e-5, (r, moved)
        torch.testing.assert_close(verts.detach(), vk - opt.param_groups[0]["lr"] * gv, rtol=1e-6, atol=1e-7)
