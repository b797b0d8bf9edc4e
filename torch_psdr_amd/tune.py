"""Autotuner for the renderer's launch shapes.

The native library exposes its run-time launch shapes through
``psdr_get_tuning`` / ``psdr_set_tuning`` (include/psdr_hip.h): the persistent
grid of the closest-hit kernel, the traversal stack entries kept in LDS, the
grids of the wavefront shading, path-replay adjoint and direct-boundary
kernels, and the lane tile of the forward render (the size of its queues).  None of them changes a result (images are bit-identical, gradients
differ only in the order of their float atomics), so the tuner may pick any of
them by time alone.

``autotune(step)`` times a caller-supplied step (for example scene rebuild +
renderD + loss + backward) under each candidate value of one parameter at a
time, keeps the fastest and repeats the sweep (coordinate descent), then leaves
the best shapes applied.  ``save`` / ``load`` keep results in a JSON file keyed
by device and workload (``workload_key``), so a training loop tunes once per
machine and workload.  The reference has no counterpart: its launch shapes
belong to OptiX and Enoki.
"""
from __future__ import annotations

import ctypes
import json
import os
from pathlib import Path
from typing import Callable, Dict, Iterable, Optional

from . import _native as N

FIELDS = ("trace_blocks_per_cu", "trace_stack_lds", "shade_blocks_per_cu", "adjoint_blocks_per_cu",
          "direct_blocks_per_cu", "tile_lanes")

DEFAULTS = {"trace_blocks_per_cu": 16, "trace_stack_lds": 16, "shade_blocks_per_cu": 8, "adjoint_blocks_per_cu": 4,
            "direct_blocks_per_cu": 8, "tile_lanes": 0}

# candidate values per parameter, swept in this order (the traversal first: it
# holds most of a step's kernel time)
SPACE: Dict[str, tuple] = {
    "trace_blocks_per_cu": (2, 4, 8, 16, 32),
    "trace_stack_lds": (8, 16, 32),
    "shade_blocks_per_cu": (2, 4, 8, 16),
    "adjoint_blocks_per_cu": (4, 8, 16, 32),
    "direct_blocks_per_cu": (2, 4, 8, 16),
    # lanes per wavefront tile of the forward render (0: every lane at once);
    # tiles bound the queue / ray / hit buffers to O(tile) lanes
    "tile_lanes": (0, 1 << 20, 1 << 21),
}


def get_tuning() -> Dict[str, int]:
    t = N.PsdrTuning()
    N.check(N.load().psdr_get_tuning(ctypes.byref(t)))
    return {f: int(getattr(t, f)) for f in FIELDS}


def set_tuning(values: Optional[Dict[str, int]] = None, **kw) -> Dict[str, int]:
    """Apply launch shapes (unnamed fields keep their current value); returns the previous ones."""
    prev = get_tuning()
    cur = dict(prev)
    for k, v in {**(values or {}), **kw}.items():
        if k not in FIELDS:
            raise KeyError(f"unknown tuning field {k!r}; fields: {FIELDS}")
        cur[k] = int(v)
    t = N.PsdrTuning(*[cur[f] for f in FIELDS])
    N.check(N.load().psdr_set_tuning(ctypes.byref(t)))
    return prev


def time_step(step: Callable[[], object], repeats: int = 3, warmup: int = 1) -> float:
    """Median wall time (ms) of ``step()`` on the current stream, synchronised per call."""
    import time

    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    ts.sort()
    return ts[len(ts) // 2]


def autotune(step: Callable[[], object], space: Optional[Dict[str, Iterable[int]]] = None, repeats: int = 3,
             rounds: int = 2, min_gain: float = 0.01, log: Optional[Callable[[str], None]] = None) -> Dict:
    """Coordinate descent over ``space`` (default SPACE) timing ``step``.

    A candidate replaces the current value only when it is faster by more than
    ``min_gain`` (relative), so timing noise does not move the shapes.  Returns
    {"tuning": best shapes (applied), "ms": their time, "start": the shapes and
    time it started from, "trials": [(shapes, ms), ...]}.
    """
    space = {k: tuple(v) for k, v in (space or SPACE).items()}
    for k in space:
        if k not in FIELDS:
            raise KeyError(f"unknown tuning field {k!r}")
    start = get_tuning()
    best = dict(start)
    set_tuning(best)
    best_ms = time_step(step, repeats=repeats, warmup=1)
    trials = [(dict(best), best_ms)]
    if log:
        log(f"autotune start {best} {best_ms:.3f} ms")
    try:
        for _ in range(max(1, rounds)):
            moved = False
            for field, values in space.items():
                for v in values:
                    if v == best[field]:
                        continue
                    cand = dict(best, **{field: v})
                    set_tuning(cand)
                    ms = time_step(step, repeats=repeats, warmup=1)
                    trials.append((cand, ms))
                    if log:
                        log(f"autotune {field}={v}: {ms:.3f} ms")
                    if ms < best_ms * (1.0 - min_gain):
                        best, best_ms, moved = cand, ms, True
                set_tuning(best)
            if not moved:
                break
    finally:
        set_tuning(best)
    return {"tuning": best, "ms": best_ms, "start": {"tuning": start, "ms": trials[0][1]}, "trials": trials}


def workload_key(scene, integrator) -> str:
    """Device + workload signature: triangles and lanes (rounded to powers of two), depth, sample counts."""
    import torch
    flat = scene.flat
    props = torch.cuda.get_device_properties(flat.vertices.device)
    cam = scene.cameras[0]
    lanes = int(integrator.cfg[1]) * int(cam.height) * int(cam.width)

    def p2(x):
        return 1 << max(0, int(x) - 1).bit_length()

    return (f"{props.name}|cu{props.multi_processor_count}|tris{p2(flat.n_tris)}|lanes{p2(lanes)}"
            f"|cfg{','.join(str(int(c)) for c in integrator.cfg)}")


def default_cache() -> Path:
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return Path(base) / "torch_psdr_amd" / "tuning.json"


def save(key: str, tuning: Dict[str, int], path: Optional[Path] = None) -> Path:
    path = Path(path) if path else default_cache()
    path.parent.mkdir(parents=True, exist_ok=True)
    data = json.loads(path.read_text()) if path.exists() else {}
    data[key] = {f: int(tuning[f]) for f in FIELDS}
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    os.replace(tmp, path)
    return path


def load(key: str, path: Optional[Path] = None, apply: bool = True) -> Optional[Dict[str, int]]:
    """Shapes stored for ``key`` (applied unless apply=False), or None."""
    path = Path(path) if path else default_cache()
    if not path.exists():
        return None
    entry = json.loads(path.read_text()).get(key)
    if entry is None:
        return None
    tuning = {f: int(entry.get(f, DEFAULTS[f])) for f in FIELDS}
    if apply:
        set_tuning(tuning)
    return tuning
