"""PyTorch custom ops over the C ABI: psdr::render and psdr::render_backward.

`psdr::render(token, cam, cfg, vertices, texels, emit, keep)` renders one
camera into a (H, W, 3) fp32 tensor on the current HIP stream; its registered
autograd formula calls `psdr::render_backward`, the hand-written adjoint
(path replay + edge sampling).  This replaces the reference's Wrap2torch
(src/python/torch_psdr.py:29-72), which bridged Enoki's tape to torch.

A token names one flattened Scene (its device buffers and host object table);
native scene handles are shared between Scene objects of the same topology, so
an optimisation loop that rebuilds its Scene every iteration re-derives the
geometry and LBVH on the GPU instead of re-creating the handle.
"""
from __future__ import annotations

import itertools
import threading
import weakref
from typing import Dict, List, Sequence, Tuple

import torch

from . import _native as N
from .flat import FlatScene

_lock = threading.Lock()
_ids = itertools.count(1)
_entries: Dict[int, Tuple["NativeScene", FlatScene]] = {}
_by_topo: Dict[tuple, "NativeScene"] = {}
_MAX_CACHED = 128
# "fused" (default) or "plain": the reference's launch shape, used as the
# measured baseline (psdr_render_forward_plain / psdr_render_backward_plain).
# Process-wide, because autograd runs the backward on its own thread.
_shape = ["fused"]


class launch_shape:
    """Context manager selecting the native launch shape of renders (and their
    backward passes) issued inside it."""

    def __init__(self, name: str):
        if name not in ("fused", "plain"):
            raise ValueError(f"unknown launch shape {name!r}")
        self.name = name

    def __enter__(self):
        with _lock:
            self.prev = _shape[0]
            _shape[0] = self.name
        return self

    def __exit__(self, *exc):
        with _lock:
            _shape[0] = self.prev
        return False


def _plain() -> bool:
    return _shape[0] == "plain"


# Native scenes used while a HIP graph is being captured (torch_psdr_amd.graph):
# one dict per open capture, NativeScene -> {"flat": the binding made before
# the capture, when the capture first used the scene without binding it anew}.
_capture_sets: List[dict] = []


def _capturing() -> bool:
    return torch.cuda.is_available() and torch.cuda.is_current_stream_capturing()


def _capture_record(ns: "NativeScene"):
    if not (_capture_sets and _capturing()):
        return None
    with _lock:
        return _capture_sets[-1].setdefault(ns, {})


class NativeScene:
    """Owns a psdr_scene handle bound to one FlatScene at a time."""

    def __init__(self, flat: FlatScene):
        self.lib = N.load()
        self.keep_alive: list = []
        self.handle = N.C.c_void_p()
        self._bound = None
        desc = flat.desc(self.keep_alive)
        N.check(self.lib.psdr_scene_create(N.C.byref(desc), N.C.byref(self.handle), N.current_stream()))
        self._hold(flat)

    def _hold(self, flat: FlatScene) -> None:
        # The native scene reads the bound buffers until the next bind: keep
        # their storage alive, detached (an autograd graph kept alive here would
        # tie later backward passes to the stream of the step that built it).
        self.buffers = tuple(t.detach() for t in (flat.vertices, flat.texels, flat.light_emit))
        self._bound = weakref.ref(flat)

    def bound_to(self, flat: FlatScene) -> bool:
        return self._bound is not None and self._bound() is flat

    def stale(self) -> None:
        """Derive the scene again at its next bind (its derived data or forward
        states were rewritten behind the host's back, e.g. by a graph replay)."""
        self._bound = None

    def bind(self, flat: FlatScene) -> None:
        rec = _capture_record(self)
        if self.bound_to(flat):
            if rec is not None and not rec:
                # a graph rendering a scene bound before the capture: every
                # replay derives that binding again first (CapturedStep)
                if any(t.grad_fn is not None for t in (flat.vertices, flat.texels, flat.light_emit)):
                    raise RuntimeError("a captured step must build its Scene itself: this Scene was built before "
                                       "the capture, with an autograd graph the captured backward cannot join")
                rec["flat"] = flat
            return
        if rec is not None:
            rec.setdefault("updated", True)
        keep_alive: list = []
        desc = flat.desc(keep_alive)
        N.check(self.lib.psdr_scene_update(self.handle, N.C.byref(desc), N.current_stream()))
        self.keep_alive = keep_alive
        self._hold(flat)

    def join(self) -> None:
        """Order the scene's open auxiliary-stream work before later work on the current stream."""
        N.check(self.lib.psdr_scene_join(self.handle, N.current_stream()))

    def info(self):
        nt, ne, dep = N.C.c_int32(), N.C.c_int32(), N.C.c_int32()
        N.check(self.lib.psdr_scene_info(self.handle, N.C.byref(nt), N.C.byref(ne), N.C.byref(dep)))
        return nt.value, ne.value, dep.value

    def __del__(self):
        try:
            if self.handle:
                self.lib.psdr_scene_destroy(self.handle)
        except Exception:
            pass


def register(flat: FlatScene) -> int:
    """Register a flattened scene; returns its token."""
    with _lock:
        ns = _by_topo.get(flat.topo_key)
    if ns is None:
        if _capturing():
            raise RuntimeError("a scene topology seen for the first time cannot be set up under HIP graph capture: "
                               "run the step eagerly before capturing it")
        ns = NativeScene(flat)
        with _lock:
            if len(_by_topo) >= _MAX_CACHED:
                _by_topo.pop(next(iter(_by_topo)))
            _by_topo[flat.topo_key] = ns
    else:
        ns.bind(flat)
    tok = next(_ids)
    with _lock:
        _entries[tok] = (ns, flat)
    return tok


def share_workspace(scenes) -> None:
    """Make the native handles of `scenes` (api.Scene objects) share one set of
    per-render buffers (psdr_scene_share_workspace): scenes rendered one after
    another then hold one set sized for the largest render, not one each."""
    handles = []
    for sc in scenes:
        ns, _ = entry(sc.token)
        if all(ns is not h for h in handles):
            handles.append(ns)
    for ns in handles[1:]:
        N.check(ns.lib.psdr_scene_share_workspace(ns.handle, handles[0].handle, N.current_stream()))


def unregister(tok: int) -> None:
    with _lock:
        _entries.pop(tok, None)


def entry(tok: int) -> Tuple[NativeScene, FlatScene]:
    return _entries[tok]


def make_config(cfg: Sequence[int]) -> N.PsdrConfig:
    c = N.PsdrConfig()
    (c.n_pass, c.spp_interior, c.enable_light_visable, c.spp_primary_edge, c.spp_secondary_edge, c.max_bounce,
     c.mis_light_samples, c.mis_bsdf_samples, c.primary_edge_preprocess_rounds, c.secondary_edge_preprocess_rounds,
     c.primary_edge_hypercube_resolution) = [int(x) for x in cfg[:11]]
    for i in range(3):
        c.secondary_edge_hypercube_resolution[i] = int(cfg[11 + i])
    return c


def make_camera(cam: Sequence[float]) -> N.PsdrCamera:
    """cam = look_from(3), look_at(3), up(3), vfov or view_width, height, width[, kind]."""
    c = N.PsdrCamera()
    for i in range(3):
        c.look_from[i] = float(cam[i])
        c.look_at[i] = float(cam[3 + i])
        c.up[i] = float(cam[6 + i])
    c.kind = int(cam[12]) if len(cam) > 12 else N.PSDR_CAMERA_PERSPECTIVE
    if c.kind == N.PSDR_CAMERA_ORTHOGRAPHIC:
        c.view_width = float(cam[9])
    else:
        c.vfov = float(cam[9])
    c.height = int(cam[10])
    c.width = int(cam[11])
    return c


def forward_native(ns: NativeScene, flat: FlatScene, cam: Sequence[float], cfg: Sequence[int],
                   keep: bool) -> torch.Tensor:
    ns.bind(flat)
    H, W = int(cam[10]), int(cam[11])
    out = torch.empty((H, W, 3), dtype=torch.float32, device=flat.vertices.device)
    c, k = make_config(cfg), make_camera(cam)
    fn = ns.lib.psdr_render_forward_plain if _plain() else ns.lib.psdr_render_forward
    N.check(fn(ns.handle, N.C.byref(c), N.C.byref(k), out.data_ptr(), int(keep), N.current_stream()))
    return out


def forward_native_multi(ns: NativeScene, flat: FlatScene, cams: Sequence[Sequence[float]], cfg: Sequence[int],
                         keep: bool) -> List[torch.Tensor]:
    """The images of up to N.PSDR_MAX_CAMERAS cameras of one scene, rendered in
    one wavefront (psdr_render_forward_multi)."""
    ns.bind(flat)
    outs = [torch.empty((int(c[10]), int(c[11]), 3), dtype=torch.float32, device=flat.vertices.device) for c in cams]
    c = make_config(cfg)
    ks = (N.PsdrCamera * len(cams))(*[make_camera(cam) for cam in cams])
    ptrs = (N.C.c_void_p * len(cams))(*[o.data_ptr() for o in outs])
    N.check(ns.lib.psdr_render_forward_multi(ns.handle, N.C.byref(c), ks, len(cams), ptrs, int(keep),
                                             N.current_stream()))
    return outs


@torch.library.custom_op("psdr::render", mutates_args=())
def render(token: int, cam: List[float], cfg: List[int], vertices: torch.Tensor, texels: torch.Tensor,
           emit: torch.Tensor, camera: torch.Tensor, keep: bool) -> torch.Tensor:
    """`camera` is the camera's parameter vector (PSDR_CAMERA_GRAD floats, the
    values of `cam[:10]`): an input only so that autograd routes its gradient."""
    ns, flat = entry(token)
    return forward_native(ns, flat, cam, cfg, keep)


@render.register_fake
def _render_fake(token, cam, cfg, vertices, texels, emit, camera, keep):
    return vertices.new_empty((int(cam[10]), int(cam[11]), 3))


@torch.library.custom_op("psdr::render_backward", mutates_args=())
def render_backward(token: int, cam: List[float], cfg: List[int], grad: torch.Tensor, vertices: torch.Tensor,
                    texels: torch.Tensor, emit: torch.Tensor, want_camera: bool) -> List[torch.Tensor]:
    ns, flat = entry(token)
    ns.bind(flat)
    c, k = make_config(cfg), make_camera(cam)
    gv = torch.zeros_like(vertices)
    gt = torch.zeros_like(texels)
    ge = torch.zeros((max(int(emit.shape[0]), 1), 3), dtype=torch.float32, device=emit.device)
    gc = torch.zeros(N.PSDR_CAMERA_GRAD, dtype=torch.float32, device=vertices.device)
    g = grad.contiguous().to(torch.float32)
    fn = ns.lib.psdr_render_backward_plain if _plain() else ns.lib.psdr_render_backward
    args = (ns.handle, N.C.byref(c), N.C.byref(k), g.data_ptr(), gv.data_ptr(), gt.data_ptr(), ge.data_ptr(),
            gc.data_ptr() if want_camera else None, N.PSDR_TERM_ALL, N.current_stream())
    rc = fn(*args)
    if rc == N.PSDR_ERR_STATE:
        # the scene keeps the forward state of its last few (camera, config)
        # pairs; render this one again when it has been evicted
        forward_native(ns, flat, cam, cfg, True)
        rc = fn(*args)
    N.check(rc)
    return [gv, gt, ge[: emit.shape[0]].clone(), gc]


@render_backward.register_fake
def _render_backward_fake(token, cam, cfg, grad, vertices, texels, emit, want_camera):
    return [torch.empty_like(vertices), torch.empty_like(texels), torch.empty_like(emit),
            vertices.new_empty((N.PSDR_CAMERA_GRAD,))]


def _setup(ctx, inputs, output):
    token, cam, cfg, vertices, texels, emit, camera, keep = inputs
    ctx.token, ctx.cam, ctx.cfg = token, cam, cfg
    ctx.entry = entry(token)  # keeps the scene alive until backward
    ctx.camera_meta = (camera.device, camera.dtype)
    ctx.save_for_backward(vertices, texels, emit)


def _backward(ctx, grad):
    vertices, texels, emit = ctx.saved_tensors
    # the Scene may be gone by now: its entry is visible to the op for this call only
    with _lock:
        added = ctx.token not in _entries
        if added:
            _entries[ctx.token] = ctx.entry
    want_camera = bool(ctx.needs_input_grad[6])
    try:
        gv, gt, ge, gc = render_backward(ctx.token, ctx.cam, ctx.cfg, grad, vertices, texels, emit, want_camera)
    finally:
        if added:
            with _lock:
                _entries.pop(ctx.token, None)
    gcam = gc.to(device=ctx.camera_meta[0], dtype=ctx.camera_meta[1]) if want_camera else None
    return None, None, None, gv, gt, ge, gcam, None


render.register_autograd(_backward, setup_context=_setup)


def pack_cams(cams: Sequence[Sequence[float]]) -> List[float]:
    """Camera parameter lists as one list: each one preceded by its length."""
    out: List[float] = []
    for c in cams:
        out.append(float(len(c)))
        out.extend(float(x) for x in c)
    return out


def unpack_cams(packed: Sequence[float]) -> List[List[float]]:
    cams, i = [], 0
    while i < len(packed):
        n = int(packed[i])
        cams.append(list(packed[i + 1:i + 1 + n]))
        i += 1 + n
    return cams


@torch.library.custom_op("psdr::render_multi", mutates_args=())
def render_multi(token: int, cams: List[float], cfg: List[int], vertices: torch.Tensor, texels: torch.Tensor,
                 emit: torch.Tensor, cameras: List[torch.Tensor], keep: bool) -> List[torch.Tensor]:
    """Several cameras of one scene in one wavefront.  `cams`: their parameter
    lists packed by pack_cams; `cameras`: their parameter vectors (inputs only
    so that autograd routes their gradients)."""
    ns, flat = entry(token)
    return forward_native_multi(ns, flat, unpack_cams(cams), cfg, keep)


@render_multi.register_fake
def _render_multi_fake(token, cams, cfg, vertices, texels, emit, cameras, keep):
    return [vertices.new_empty((int(c[10]), int(c[11]), 3)) for c in unpack_cams(cams)]


def _setup_multi(ctx, inputs, output):
    token, cams, cfg, vertices, texels, emit, cameras, keep = inputs
    ctx.token, ctx.cams, ctx.cfg = token, unpack_cams(cams), cfg
    ctx.entry = entry(token)
    ctx.camera_meta = [(c.device, c.dtype) for c in cameras]
    ctx.save_for_backward(vertices, texels, emit)


def _backward_multi(ctx, grads):
    """Each camera's backward finds its own forward state; gradients add up."""
    vertices, texels, emit = ctx.saved_tensors
    with _lock:
        added = ctx.token not in _entries
        if added:
            _entries[ctx.token] = ctx.entry
    cam_needs = ctx.needs_input_grad[6]
    gv = gt = ge = None
    gcams = []
    try:
        for k, (cam, g) in enumerate(zip(ctx.cams, grads)):
            want_camera = bool(cam_needs[k]) if isinstance(cam_needs, (list, tuple)) else bool(cam_needs)
            if g is None:
                gcams.append(None)
                continue
            v, t, e, c = render_backward(ctx.token, cam, ctx.cfg, g, vertices, texels, emit, want_camera)
            gv = v if gv is None else gv + v
            gt = t if gt is None else gt + t
            ge = e if ge is None else ge + e
            dev, dt = ctx.camera_meta[k]
            gcams.append(c.to(device=dev, dtype=dt) if want_camera else None)
    finally:
        if added:
            with _lock:
                _entries.pop(ctx.token, None)
    return None, None, None, gv, gt, ge, gcams, None


render_multi.register_autograd(_backward_multi, setup_context=_setup_multi)
