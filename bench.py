"""Benchmark: fused forward + adjoint step of the differentiable path tracer on one MI355X.

Workload (default, the 512x512 / 16 spp / depth-3 vertex + texture gradient step):
icosphere level 5 (20480 tris, vertices radially perturbed by N(0, 0.02), seed 3)
with a 512x512 RGB texture, 10x10 ground quad, 2x2 quad light (radiance 10);
camera (0,1,5) -> origin, vfov 45 deg; 512x512, 16 spp, max_bounce 3,
1 light + 1 BSDF MIS sample, spp_primary_edge = spp_secondary_edge = 4 (1M
boundary samples each); L2 loss against the unperturbed mesh rendered with a
second texture.  A step = Scene rebuild (triangle records, edges, samplers,
LBVH) + renderD forward + loss + backward (interior path replay, primary and
direct boundary terms) w.r.t. vertices and texels; synthetic data.

The step runs eagerly (warm-up, the timed steps with the traversal kernel's
launches timed for the roofline) and is then captured into a HIP graph
(torch_psdr_amd.graph.CapturedStep) whose replays are timed the same way; by
default the replay time is the reported step time, the eager one is given
beside it.  Before the timed region the autotuner (torch_psdr_amd.tune) picks
the launch shapes by timing this step (reported under "autotune"; --no-autotune
keeps the defaults).

After the timed regions: the same step in the reference's launch shape (the
plain GPU path), the gradient error of both shapes against directional central
differences of the loss (grad_error), one step with the traversal's counting
instance (traversal_counters) and the CPU oracle on a bounded sample.

Prints one JSON line.  value = Mrays/s of reference-equivalent rays:
R*N (N = spp*H*W, R = 1 + D*(L+B+1) = 10) + 20 rays per primary-boundary sample
+ 3 per direct-boundary sample, divided by the step time.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the default launch shapes instead of the autotuner's choice")
    ap.add_argument("--tuning", default=None,
                    help="launch shapes to time instead of tuning, e.g. trace_blocks_per_cu=16,direct_blocks_per_cu=16")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="steps of the reference-shaped baseline path to time (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU-baseline sample time")
    ap.add_argument("--graph", choices=("value", "report", "off"), default="value",
                    help="the step captured into a HIP graph and replayed: reported as the value (the eager "
                         "step's time beside it), reported beside the eager value, or not run")
    ap.add_argument("--workload", choices=("c2", "c0", "c1", "c3", "c4", "trace"), default="c2",
                    help="c2: the 512^2 / 16 spp vertex + texture step (default); c0: the 256^2 / 4 spp forward "
                         "render; c1: the 256^2 / 8 spp texture step; c3: 64 scenes of uneven size, rendered and "
                         "differentiated in turn on shared buffers; c4: the 1024^2 / 64 spp terrain step; trace: "
                         "the closest-hit traversal alone on the c2 scene (camera, diffuse-bounce, toward-light and "
                         "random rays)")
    ap.add_argument("--no-grad-error", action="store_true",
                    help="skip the gradient-error leg (directional central differences of the loss)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scales rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 28, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch
    import torch_psdr_amd as dr
    import torch_psdr_amd._native as N
    from torch_psdr_amd import ops
    from torch_psdr_amd import workloads as W

    dev = torch.device("cuda:0")
    if args.workload == "c3":
        bench_c3(args, torch, dr, W, dev)
        return
    if args.workload == "trace":
        bench_trace(args, torch, dr, W, N, ops, dev)
        return
    if args.workload in ("c0", "c1", "c4"):
        bench_config(args, torch, dr, W, dev)
        return
    parts = W.c2_parts(dr, dev, res=args.res, spp=args.spp)
    integ = parts["integrator"]
    with torch.no_grad():
        target = integ.renderC(parts["scene"](parts["V0"], parts["tex_target"]))[0]
    verts = parts["Vp"].clone().requires_grad_(True)
    tex = parts["tex_opt"].clone().requires_grad_(True)

    last = {}

    def step():
        verts.grad = None
        tex.grad = None
        scene = parts["scene"](verts, tex)
        # the native handle only: holding the Scene would keep this step's
        # autograd graph alive into the graph capture below
        last["ns"] = ops.entry(scene.token)[0]
        img = integ.renderD(scene)[0]
        loss = ((img - target) ** 2).mean()
        loss.backward()
        # the caller keeps only the value: a live autograd graph would tie the
        # parameters' gradient accumulators to this step's stream
        return loss.detach()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    tuned = None
    if args.tuning:
        from torch_psdr_amd import tune
        tune.set_tuning({k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))})
        tuned = {"tuning": tune.get_tuning(), "given": True}
    elif not args.no_autotune:
        # launch shapes (traversal grid and LDS stack depth, shading / adjoint /
        # boundary grids) chosen by timing this step, before the timed region
        from torch_psdr_amd import tune
        t_tune = time.perf_counter()
        res_t = tune.autotune(step, repeats=3, rounds=2)
        tuned = {"tuning": res_t["tuning"], "default": res_t["start"]["tuning"],
                 "default_ms": round(res_t["start"]["ms"], 3), "tuned_ms": round(res_t["ms"], 3),
                 "trials": len(res_t["trials"]), "seconds": round(time.perf_counter() - t_tune, 1)}
        print(f"autotune: {tuned}", file=sys.stderr, flush=True)
    # traversal launches are bracketed by HIP events on the library's stream
    N.check(N.load().psdr_trace_timing(1, None, None, None, None))
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record()
    for _ in range(args.steps):
        loss = step()
    t1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    tr_ms, tr_launch, tr_bytes, tr_busy = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
    N.check(N.load().psdr_trace_timing(0, ctypes.byref(tr_ms), ctypes.byref(tr_launch), ctypes.byref(tr_bytes),
                                       ctypes.byref(tr_busy)))
    ms = t0.elapsed_time(t1) / args.steps
    ms = max(ms, wall * 1e3 / args.steps)

    res, spp, spe = args.res, args.spp, parts["spp_edge"]
    D, L, B = 3, 1, 1
    R = 1 + D * (L + B + 1)
    N_int = spp * res * res
    N_e = spe * res * res
    rays = R * N_int + 20 * N_e + 3 * N_e
    mrays = rays / (ms * 1e-3) / 1e6

    from torch_psdr_amd import tune
    roof = trace_roofline(tr_busy.value, tr_ms.value, tr_launch.value, tr_bytes.value,
                          default_workload=(res, spp) == (512, 16), shape=tune.get_tuning())
    # rays the interior render actually traced (after compaction; the last depth's
    # extension is not traced), beside the reference-equivalent R*N
    ns = last.pop("ns")
    counts = (ctypes.c_int64 * (D + 2))()
    traced = ctypes.c_int64()
    N.check(N.load().psdr_last_render_stats(ns.handle, counts, D + 2, ctypes.byref(traced)))
    interior_rays = {"reference_equivalent": R * N_int, "traced": int(traced.value),
                     "paths_per_depth": [int(c) for c in counts]}
    graph = None
    if args.graph != "off":
        graph = graph_step_timing(torch, step, args.steps, rays)
        graph["eager_ms_per_step"] = round(ms, 3)
        if args.graph == "value":
            ms, mrays = graph["ms_per_step"], graph["value"]
    plain = None
    grads = {}
    if args.plain_steps > 0:
        plain = plain_baseline(torch, ops, step, verts, tex, args.plain_steps, rays, grads)
        plain["speedup"] = round(plain["ms_per_step"] / ms, 2)
    if "fused" not in grads and not args.no_grad_error:
        step()
        torch.cuda.synchronize()
        grads["fused"] = (verts.grad.clone(), tex.grad.clone())
    gerr = None if args.no_grad_error else grad_error(torch, parts, integ, target, verts, tex, grads)
    counters = trace_counter_pass(torch, N, step)
    step_bytes = step_algorithmic_bytes(N_int, N_e, spp, D, L, B)
    cpu = None if args.no_cpu_baseline else cpu_baseline(parts, spp, args.cpu_seconds)
    out = {
        "metric": "fwd+bwd ms/iter & Mrays/s at 512^2 16spp vs plain GPU path, equal grad error",
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"vertex+texture gradient step, icosphere L5 (20480 tris) + ground + light, "
                               f"{res}x{res}, {spp} spp, depth 3, edge spp {spe}",
                   "rays_per_step": rays, "interior_rays": interior_rays, "loss": float(loss.detach().cpu()),
                   "grad_abs_max": {"vertices": float(verts.grad.abs().max()), "texels": float(tex.grad.abs().max())}},
        "roofline": roof,
        "step_roofline": {"bound": "hbm", "achieved": round(step_bytes / (ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                          "algorithmic_bytes": step_bytes},
        "cpu_baseline": cpu,
        "plain_gpu_path": plain,
        "grad_error": gerr,
        "traversal_counters": counters,
        "graph_step": graph,
        "autotune": tuned,
    }
    print(json.dumps(out))


def bench_c3(args, torch, dr, W, dev):
    """64 scenes of uneven size (geodesic spheres of 980-184,320 triangles,
    128^2-1024^2, 4-64 spp, depth 3, one edge sample per pixel for each boundary
    term): one step renders and differentiates every scene in turn (vertex +
    texel gradients), all scenes sharing one set of per-render buffers."""
    t_setup = time.perf_counter()
    parts = W.c3_parts(dr, dev)
    print(f"c3: {len(parts)} scenes set up in {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    def scene_step(b):
        b.vertices.grad = None
        b.tex.grad = None
        img = b.integrator.renderD(b.scene(b.vertices, b.tex))[0]
        ((img - b.target) ** 2).mean().backward()

    def step():  # every scene in turn on one stream
        W.c3_step(parts, scene_step)

    D, L, B = 3, 1, 1
    R = 1 + D * (L + B + 1)
    lanes = sum(b.res * b.res * b.spp for b in parts)
    edge = sum(b.res * b.res for b in parts)
    rays = R * lanes + 23 * edge
    steps = max(1, min(args.steps, 3))
    for _ in range(max(1, min(args.warmup, 1))):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms_eager = (time.perf_counter() - t0) * 1e3 / steps
    print(f"c3: eager {ms_eager:.1f} ms per step", file=sys.stderr, flush=True)
    graph = None
    if args.graph != "off":
        graph = graph_step_timing(torch, step, steps, rays)
        graph["eager_ms_per_step"] = round(ms_eager, 3)
    forms = {"eager": ms_eager}
    if graph is not None and args.graph == "value":
        forms["graph replay"] = graph["ms_per_step"]
    form, ms = min(forms.items(), key=lambda kv: kv[1])
    print(json.dumps({
        "metric": "fwd+bwd Mrays/s, 64 scenes of uneven size in turn on shared buffers",
        "value": round(rays / (ms * 1e-3) / 1e6, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 geodesic spheres (seed 4), 980-184320 tris, 128^2-1024^2, 4/16/64 spp, depth 3, "
                               "edge spp 1, vertex + texel gradients",
                   "scenes": len(parts), "lanes": lanes, "boundary_samples_per_term": edge, "rays_per_step": rays,
                   "tris": [b.n_tris for b in parts], "res": [b.res for b in parts], "spp": [b.spp for b in parts],
                   "form": form},
        "forms_ms_per_step": {k: round(v, 3) for k, v in forms.items()},
        "graph_step": graph}))


def bench_config(args, torch, dr, W, dev):
    """The other configurations, each timed as --steps eager steps after
    --warmup (synchronised around the timed region):
    c0: icosphere L3 + ground + light, 256^2, 4 spp, depth 2, renderC;
    c1: the same scene with a 256^2 texture (seed 1) on the sphere, 256^2, 8 spp,
        depth 2, renderD + L2 loss against the seed-2 texture's image + backward
        (texel gradients);
    c4: 1024^2 fBm terrain (2.09 M triangles, two lights, 1024^2 texture),
        1024^2, 64 spp, depth 4, renderD + L2 loss + backward (height and texel
        gradients)."""
    D, L, B = 2, 1, 1
    spp_edge = 0
    if args.workload == "c0":
        scene, integ, _ = W.build_scene(dr, dev, level=3, res=256, spp=4, max_bounce=2)
        res, spp = 256, 4

        def step():
            integ.renderC(scene)
    elif args.workload == "c1":
        tex = W.texture(1, 256, dev).requires_grad_(True)
        scene, integ, _ = W.build_scene(dr, dev, level=3, res=256, tex=tex, spp=8, max_bounce=2)
        with torch.no_grad():
            tgt_scene, _, _ = W.build_scene(dr, dev, level=3, res=256, tex=W.texture(2, 256, dev), spp=8,
                                            max_bounce=2)
            target = integ.renderC(tgt_scene)[0]
        res, spp, spp_edge = 256, 8, 1

        def step():
            tex.grad = None
            img = integ.renderD(scene)[0]
            ((img - target) ** 2).mean().backward()
    else:
        p = W.c4_parts(dr, dev)
        h = p["heights"].clone().requires_grad_(True)
        tex = p["tex"].clone().requires_grad_(True)
        integ = p["integrator"]
        res, spp, spp_edge, D = 1024, 64, 1, 4

        def step():
            h.grad = None
            tex.grad = None
            img = integ.renderD(p["scene"](h, tex))[0]
            ((img - 0.5) ** 2).mean().backward()
    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    steps = max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    R = 1 + D * (L + B + 1)
    n = res * res * spp
    n_e = res * res * spp_edge
    rays = R * n + 23 * n_e
    names = {"c0": "forward render, icosphere L3 (1280 tris) + ground + light, 256x256, 4 spp, depth 2",
             "c1": "texture-gradient step, icosphere L3 + 256^2 texture, 256x256, 8 spp, depth 2, edge spp 1",
             "c4": "height + texture gradient step, 1024^2 fBm terrain (2.09M tris), two lights, 1024x1024, "
                   "64 spp, depth 4, edge spp 1"}
    print(json.dumps({"metric": f"{'fwd' if args.workload == 'c0' else 'fwd+bwd'} ms/iter & Mrays/s",
                      "value": round(rays / (ms * 1e-3) / 1e6, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": steps,
                      "warmup": max(args.warmup, 1), "ms_per_step": round(ms, 3), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                      "config": {"workload": names[args.workload], "lanes": n, "boundary_samples_per_term": n_e,
                                 "rays_per_step": rays, "form": "eager"}}))


def bench_trace(args, torch, dr, W, N, ops, dev):
    """The closest-hit traversal alone (psdr_scene_trace, one launch per ray
    set) on the 512x512 / 16 spp step's scene: camera rays of 12 spp, diffuse
    bounce rays from their hits (cosine-weighted about the sphere's or the
    ground's normal), rays from those hits toward random points of the light,
    and 3 M random rays through the scene.  Median of --steps launches per set
    after one warm-up; the checksums (sum of hit ids and of hit distances) let
    two builds be compared for equal results."""
    import ctypes
    parts = W.c2_parts(dr, dev)
    scene = parts["scene"](parts["Vp"], parts["tex_opt"])
    ns, flat = ops.entry(scene.token)
    ns.bind(flat)
    lib = ns.lib
    cam = ops.make_camera(scene.cameras[0].params())
    spp = 12
    n = 512 * 512 * spp
    o = torch.empty((n, 3), device=dev)
    d = torch.empty((n, 3), device=dev)
    pix = torch.empty((n,), dtype=torch.int32, device=dev)
    N.check(lib.psdr_camera_rays(ctypes.byref(cam), spp, o.data_ptr(), d.data_ptr(), pix.data_ptr(),
                                 N.current_stream()))
    if args.tuning:
        from torch_psdr_amd import tune
        tune.set_tuning({k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))})
    out = {}

    def run(name, ro, rd):
        m = ro.shape[0]
        tri = torch.empty((m,), dtype=torch.int32, device=dev)
        t, u, v = (torch.empty((m,), device=dev) for _ in range(3))

        def f():
            N.check(lib.psdr_scene_trace(ns.handle, ro.data_ptr(), rd.data_ptr(), m, tri.data_ptr(), t.data_ptr(),
                                         u.data_ptr(), v.data_ptr(), N.current_stream()))
        f()
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(args.steps, 1)):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            f()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        med = sorted(ts)[len(ts) // 2]
        hit = tri >= 0
        out[name] = {"rays": m, "median_us": round(med * 1e3, 1), "grays_per_s": round(m / med / 1e6, 3),
                     "hit_fraction": round(float(hit.float().mean()), 4),
                     "checksum": [int(tri.long().sum()), float(t[hit].double().sum())]}
        print(f"trace {name}: {m} rays, {med * 1e3:.1f} us", file=sys.stderr, flush=True)
        return tri, t

    tri, t = run("camera", o, d)
    hit = tri >= 0
    p = o + d * t[:, None]
    g = torch.Generator(device=dev).manual_seed(5)
    up = torch.tensor([0.0, 1.0, 0.0], device=dev)
    nrm = torch.where((p[:, 1] < -0.999)[:, None], up, torch.nn.functional.normalize(p, dim=1))
    dirs = torch.nn.functional.normalize(nrm + torch.nn.functional.normalize(
        torch.randn((n, 3), device=dev, generator=g), dim=1), dim=1)
    sel = hit.nonzero().squeeze(1)
    ob = (p + nrm * 1e-3)[sel].contiguous()
    run("bounce", ob, dirs[sel].contiguous())
    lp = torch.rand((sel.shape[0], 3), device=dev, generator=g) * 2 - 1
    lp[:, 1] = 4.0
    run("tolight", ob, torch.nn.functional.normalize(lp - ob, dim=1).contiguous())
    gc = torch.Generator().manual_seed(1)
    orr = (torch.rand((3_000_000, 3), generator=gc) * 2 - 1) * torch.tensor([6.0, 3.0, 6.0]) + \
        torch.tensor([0.0, 2.0, 0.0])
    tg = (torch.rand((3_000_000, 3), generator=gc) * 2 - 1) * 1.5
    run("random", orr.to(dev).contiguous(), torch.nn.functional.normalize(tg - orr, dim=1).to(dev).contiguous())
    print(json.dumps({"metric": "closest-hit traversal Grays/s per ray set (c2 scene)", "unit": "Grays/s",
                      "n_gpus": 1, "steps": args.steps, "data": "synthetic",
                      "config": {"workload": "closest-hit traversal, icosphere L5 (20480 tris) + ground + light",
                                 "tuning": args.tuning}, "sets": out}))


def graph_step_timing(torch, step, n_steps, rays):
    """The same step (Scene rebuild with its LBVH, renderD, loss, backward)
    captured once into a HIP graph and replayed n_steps times."""
    from torch_psdr_amd.graph import CapturedStep
    print("graph step: warm-up and capture", file=sys.stderr, flush=True)
    cs = CapturedStep(step, warmup=1)
    print("graph step: first replay", file=sys.stderr, flush=True)
    cs.replay()
    torch.cuda.synchronize()
    print("graph step: timed replays", file=sys.stderr, flush=True)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record()
    for _ in range(n_steps):
        cs.replay()
    t1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = max(t0.elapsed_time(t1) / n_steps, wall * 1e3 / n_steps)
    return {"ms_per_step": round(ms, 3), "value": round(rays / (ms * 1e-3) / 1e6, 3), "unit": "Mrays/s",
            "steps": n_steps}


def trace_counter_pass(torch, N, step):
    """One more eager step with the traversal's counting instance (outside the
    timed region): rays traced, BVH nodes fetched and triangles tested per ray,
    and the closest-hit rays whose hit is a near tie (closest and second-closest
    triangle within 4 ulp) or grazes an edge (barycentrics within 1e-6), the
    rays whose identity another implementation's arithmetic could change."""
    lib = N.load()
    N.check(lib.psdr_trace_counters(1, None, 0))
    step()
    torch.cuda.synchronize()
    c = (ctypes.c_int64 * N.PSDR_TRACE_COUNTERS)()
    N.check(lib.psdr_trace_counters(0, c, N.PSDR_TRACE_COUNTERS))
    rays = max(int(c[0]), 1)
    return {"rays": int(c[0]), "nodes_per_ray": round(c[1] / rays, 3), "triangles_per_ray": round(c[2] / rays, 3),
            "exact_tests_per_ray": round(c[3] / rays, 3), "hits": int(c[6]), "closest_hit_rays": int(c[7]),
            "near_ties": int(c[4]), "grazing": int(c[5])}


def grad_error(torch, parts, integ, target, verts, tex, grads, n_dirs=3, h_tex=1e-2, h_vert=1e-4, seed=17):
    """Gradient error against central differences of the same estimator (same
    sample streams): along n_dirs random directions w per parameter block, the
    relative L2 error over the directions of dL.w (renderD's gradient) against
    (L(p + h w) - L(p - h w)) / 2h with L the step's L2 loss rendered by renderC,
    in double.  Texel directions are N(0, 1) per texel channel; vertex
    directions N(0, 1) per coordinate.  The texel part of the estimator is
    smooth in the texels, so its error is the estimator's own; the vertex part
    carries the reference's boundary normalisation (samples summed per pixel
    over the edge spp only, hypercube cell pmf used as a density), which a
    difference quotient of the image does not, so its error is large by
    construction and equal for both launch shapes."""
    g = torch.Generator().manual_seed(seed)

    def loss(v, t):
        with torch.no_grad():
            img = integ.renderC(parts["scene"](v, t))[0]
            return float(((img.double() - target.double()) ** 2).mean())

    out = {}
    for name, p, h in (("texels", tex, h_tex), ("vertices", verts, h_vert)):
        dirs = [torch.randn(p.shape, generator=g).to(p.device) for _ in range(n_dirs)]
        fd = []
        for w in dirs:
            plus = loss(verts.detach() + h * w, tex.detach()) if name == "vertices" else loss(verts.detach(), tex.detach() + h * w)
            minus = loss(verts.detach() - h * w, tex.detach()) if name == "vertices" else loss(verts.detach(), tex.detach() - h * w)
            fd.append((plus - minus) / (2 * h))
        fd_n = sum(x * x for x in fd) ** 0.5
        rec = {"h": h, "directions": n_dirs, "fd": [round(x, 9) for x in fd]}
        for shape, gg in grads.items():
            gp = gg[0] if name == "vertices" else gg[1]
            an = [float((gp.double() * w.double()).sum()) for w in dirs]
            err = sum((a - b) ** 2 for a, b in zip(an, fd)) ** 0.5 / max(fd_n, 1e-300)
            rec[shape] = {"rel_l2_error": round(err, 6), "directional": [round(x, 9) for x in an]}
        out[name] = rec
    return out


def plain_baseline(torch, ops, step, verts, tex, n_steps, rays, grads):
    """The same step in the reference's launch shape (full-width launches without
    compaction, one synchronised trace launch per ray kind, boundary samples in
    lane order, gradient terms one after another).  Both shapes compute the same
    estimator from the same samples, so their gradient error is equal; the
    largest gradient difference (accumulation order only) is reported."""
    step()
    torch.cuda.synchronize()
    g_fused = (verts.grad.clone(), tex.grad.clone())
    grads["fused"] = g_fused
    with ops.launch_shape("plain"):
        step()  # warm-up, also the comparison step
        torch.cuda.synchronize()
        g_plain = (verts.grad.clone(), tex.grad.clone())
        grads["plain"] = g_plain
        t0 = time.perf_counter()
        for _ in range(n_steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / n_steps
    diff = max(float((a - b).abs().max() / b.abs().max().clamp_min(1e-30)) for a, b in zip(g_plain, g_fused))
    return {"ms_per_step": round(ms, 3), "value": round(rays / (ms * 1e-3) / 1e6, 3), "unit": "Mrays/s",
            "steps": n_steps, "grad_max_rel_diff_vs_fused": diff}


# HBM bytes per k_trace launch on the default workload from two rocprofv3 --pmc
# runs (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, summed over one step's 11
# traversal launches), each scaled by what the counter reads for a 1 GiB
# device-to-device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0), at every launch
# shape the autotuner can pick (tune.SPACE); another shape reports null.
# Commands and per-launch table: profiles/k_trace_gradient_step_512px_16spp.md.
TRACE_TRAFFIC = [
    {"shape": {"trace_blocks_per_cu": 2, "trace_stack_lds": 8}, "bytes_per_launch": 76078333},
    {"shape": {"trace_blocks_per_cu": 2, "trace_stack_lds": 16}, "bytes_per_launch": 75866333},
    {"shape": {"trace_blocks_per_cu": 2, "trace_stack_lds": 32}, "bytes_per_launch": 75945836},
    {"shape": {"trace_blocks_per_cu": 4, "trace_stack_lds": 8}, "bytes_per_launch": 75486316},
    {"shape": {"trace_blocks_per_cu": 4, "trace_stack_lds": 16}, "bytes_per_launch": 74916919},
    {"shape": {"trace_blocks_per_cu": 4, "trace_stack_lds": 32}, "bytes_per_launch": 75034810},
    {"shape": {"trace_blocks_per_cu": 8, "trace_stack_lds": 8}, "bytes_per_launch": 75108820},
    {"shape": {"trace_blocks_per_cu": 8, "trace_stack_lds": 16}, "bytes_per_launch": 74892576},
    {"shape": {"trace_blocks_per_cu": 8, "trace_stack_lds": 32}, "bytes_per_launch": 74995843},
    {"shape": {"trace_blocks_per_cu": 16, "trace_stack_lds": 8}, "bytes_per_launch": 75082877},
    {"shape": {"trace_blocks_per_cu": 16, "trace_stack_lds": 16}, "bytes_per_launch": 74857536},
    {"shape": {"trace_blocks_per_cu": 16, "trace_stack_lds": 32}, "bytes_per_launch": 75008445},
    {"shape": {"trace_blocks_per_cu": 32, "trace_stack_lds": 8}, "bytes_per_launch": 75659017},
    {"shape": {"trace_blocks_per_cu": 32, "trace_stack_lds": 16}, "bytes_per_launch": 75488189},
    {"shape": {"trace_blocks_per_cu": 32, "trace_stack_lds": 32}, "bytes_per_launch": 76134781},
]


def trace_roofline(busy_ms, event_ms, launches, bytes_, default_workload, shape):
    """Dominant kernel (closest-hit LBVH traversal): its launches inside the
    timed region.  achieved = algorithmic bytes (24 B ray read + 12 B hit
    written per ray) / summed execution time, each launch's execution time
    being its latest workgroup end minus its earliest workgroup start on the
    device clock (what a kernel trace reports as the dispatch's duration).
    The HIP-event durations of the same launches are given beside it: they
    also count time a launch waits for CUs held by the boundary terms'
    streams, and understate the kernel's rate."""
    sec = busy_ms * 1e-3
    ach = bytes_ / sec / 1e9 if sec > 0 else 0.0
    ev_sec = event_ms * 1e-3
    ev_ach = bytes_ / ev_sec / 1e9 if ev_sec > 0 else 0.0
    measured = [t for t in TRACE_TRAFFIC if all(shape.get(k) == v for k, v in t["shape"].items())]
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5),
            "traffic": measured[0]["bytes_per_launch"] if (default_workload and measured) else None,
            "traffic_profiled_shapes": [t["shape"] for t in TRACE_TRAFFIC],
            "kernel": "k_trace (closest-hit LBVH traversal)", "launches": launches,
            "algorithmic_bytes_per_launch": round(bytes_ / max(launches, 1)),
            "avg_launch_ms": round(busy_ms / max(launches, 1), 4),
            "timing": "in-kernel span (earliest workgroup start to latest workgroup end)",
            "event_timed": {"avg_launch_ms": round(event_ms / max(launches, 1), 4), "achieved": round(ev_ach, 2),
                            "frac": round(ev_ach / HBM_PEAK_GBS, 5)},
            "grays_per_s": round(bytes_ / 36.0 / sec / 1e9, 3) if sec > 0 else 0.0}


def step_algorithmic_bytes(n_int, n_e, spp, D, L, B):
    """Minimal wavefront stream traffic of one step: 72 B per ray query (ray
    written + read, hit written + read), 72 B of path state per shading depth,
    12 B film per pixel; the adjoint adds dL/dI (12 B) and 48 B per depth of
    texel + vertex gradient; boundary samples cost 72 B per hit (20 hits per
    primary, 3 per direct sample)."""
    R = 1 + D * (L + B + 1)
    b_fwd = 72.0 * R + 72.0 * D + 12.0 / spp
    b_bwd = b_fwd + 12.0 + 48.0 * D
    return int(round((b_fwd + b_bwd) * n_int + 72.0 * 23 * n_e))


def host_cpus():
    """The host's CPUs as this process sees them: logical CPUs (nproc), the
    CPUs it may run on (affinity), the cgroup CPU quota if one is set, and the
    model name.  The oracle runs one thread per CPU it may use: the affinity
    count, or the quota (rounded up) when that is smaller."""
    import math
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = affinity if quota is None else max(1, min(affinity, math.ceil(quota)))
    return {"nproc": nproc, "affinity": affinity, "cgroup_cpu_quota": quota, "model": model,
            "threads_used": threads}


def cpu_baseline(parts, spp, target_s):
    """The CPU oracle (a port of the reference algorithm) timed on this host on a
    bounded sample: the same scene at a reduced resolution with the same spp,
    renderD forward + forward-mode derivative (interior + both boundary terms);
    the resolution is chosen from a 64^2 trial so the sample takes ~target_s."""
    try:
        import numpy as np
        import torch
        from oracle import binding as O
        from torch_psdr_amd.flat import flatten
        import torch_psdr_amd as dr
        from torch_psdr_amd import workloads as W
        host = host_cpus()
        cores = host["threads_used"]
        O.set_threads(cores)
        cpu = torch.device("cpu")
        V0, F = W.icosphere(5)
        Vp = W.perturbed_sphere(5, 3)
        tex = parts["tex_opt"].cpu()
        sphere = dr.Mesh(torch.tensor(Vp), torch.tensor(W.sphere_uv(V0)), torch.tensor(F), torch.tensor(F),
                         dr.DiffuseMaterial(tex))
        Vg, Fg = W.quad(10.0, -1.0)
        ground = dr.Mesh(torch.tensor(Vg), None, torch.tensor(Fg), None, dr.DiffuseMaterial(torch.tensor([0.5] * 3)))
        Vl, Fl = W.quad(2.0, 4.0, facing_down=True)
        light = dr.AreaLight(torch.tensor(Vl), torch.tensor(Fl), torch.tensor([10.0] * 3))
        flat = flatten([sphere, ground], [light], cpu)
        rng = np.random.default_rng(0)
        osc = O.Scene(flat, d_vertices=rng.normal(size=tuple(flat.vertices.shape)).astype(np.float32),
                      d_texels=rng.normal(size=flat.texels.numel()).astype(np.float32))
        cfg = [1, spp, 0, 4, 4, 3, 1, 1, 1, 1, 10000, 10000, 6, 6]

        def run(r):
            cam = [0.0, 1.0, 5.0, 0, 0, 0, 0, 1, 0, 3.14159265 / 4, r, r]
            t = time.perf_counter()
            O.render(osc, cfg, cam, mode=1, boundary=True)
            return time.perf_counter() - t

        t64 = run(64)
        sample_res = int(min(512, max(64, 64 * (target_s / max(t64, 1e-3)) ** 0.5)))
        sample_res -= sample_res % 8
        reps, sec = 0, 0.0
        while reps == 0 or (sec < target_s and reps < 8):
            sec += run(sample_res)
            reps += 1
        n_int = spp * sample_res * sample_res
        n_e = 4 * sample_res * sample_res
        rays = reps * (10 * n_int + 23 * n_e)
        return {"value": round(rays / sec / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
                "host": host,
                "sample": f"oracle renderD forward + forward-mode derivative along one direction of every vertex "
                          f"and texel (interior, primary and direct boundary; the oracle has no reverse mode) at "
                          f"{sample_res}x{sample_res}, {spp} spp, edge spp 4, same scene, {reps} run(s); "
                          f"{sec:.1f} s on {cores} threads"}
    except Exception as e:  # the baseline is reported, never required
        return {"value": None, "unit": "Mrays/s", "cores": None, "kind": "port", "sample": f"failed: {e}"}


if __name__ == "__main__":
    main()
