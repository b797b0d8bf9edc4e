"""HIP-graph capture of an optimisation step.

The reference synchronises the host after every ray cast
(src/scene_optix.cc:275) and rebuilds its Enoki graph each iteration; here a
whole step — Scene rebuild (triangle records, edges, samplers, LBVH), renderD,
loss, backward (interior path replay and both boundary terms, on their
auxiliary streams) and the optimizer update — is recorded once into a HIP graph
and replayed with one launch.  The native calls allocate nothing and never
synchronise the host once they have run eagerly at the same sizes, so the
captured step is the eager step's kernels, with the eager step's results.

    step = CapturedStep(lambda: one_step())   # warm-up runs + capture
    for it in range(n):
        loss = step.replay()                   # the tensors fn() returned

Rules a captured step keeps (each is checked where it can be):
- fn() must run eagerly once at the same sizes first (the warm-up does); a
  native buffer that would grow under capture is an error;
- fn() builds its Scene itself, as the reference's optimisation loops do (the
  Scene copies the parameters at construction); a Scene built before the
  capture is derived again before every replay, and refused when it carries an
  autograd graph;
- no autograd graph of an earlier eager step may be alive at capture (keep
  only detached values of eager losses): its gradient accumulators would tie
  the captured backward to the eager step's stream;
- camera parameters are read on the host before capture; a camera whose
  tensors change (e.g. are optimised) between eager reads and capture is an
  error;
- replays and eager renders share the scene's buffers and run in stream order;
  after a replay, the next eager use of a scene it touched derives the scene
  again (its geometry may have changed) and renders its forward state again.
"""
from __future__ import annotations

from typing import Any, Callable, Optional

import torch

from . import ops


class CapturedStep:
    """fn, run `warmup` times eagerly on a side stream and then captured into a
    torch.cuda.CUDAGraph (a HIP graph on ROCm); replay() launches the graph and
    returns what fn() returned under capture (static tensors, rewritten by every
    replay)."""

    def __init__(self, fn: Callable[[], Any], warmup: int = 2, pool: Optional[Any] = None):
        if not torch.cuda.is_available():
            raise RuntimeError("CapturedStep needs a HIP device")
        if warmup < 1:
            raise ValueError("at least one eager warm-up run is needed before capture")
        self.fn = fn
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        used: dict = {}
        with ops._lock:
            ops._capture_sets.append(used)
        try:
            with torch.cuda.graph(self.graph, pool=pool):
                self.out = fn()
        finally:
            with ops._lock:
                ops._capture_sets.remove(used)
        # the native scenes the graph points into stay alive with it; scenes
        # the graph renders with a binding made before the capture are derived
        # from that binding again before each replay
        self.scenes = list(used)
        self.prebind = [(ns, rec["flat"]) for ns, rec in used.items() if "flat" in rec]
        self._stale()

    def _stale(self) -> None:
        # the graph rewrites these scenes' derived data and forward state: the
        # next eager bind derives the scene again and drops its forward states
        for ns in self.scenes:
            ns.stale()

    def replay(self) -> Any:
        # eager boundary work still open on a scene's auxiliary streams comes first
        for ns in self.scenes:
            ns.join()
        for ns, flat in self.prebind:
            ns.bind(flat)
        self.graph.replay()
        self._stale()
        return self.out
