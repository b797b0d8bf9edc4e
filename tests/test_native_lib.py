"""The C-ABI library loads and exports every entry point include/psdr_hip.h declares (no GPU calls)."""
import ctypes
import re

import pytest
from pathlib import Path

import torch_psdr_amd._native as N

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "psdr_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psdr_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_declared_symbols():
    lib = N.load()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(N.EXPORTED)


def test_abi_version_and_error_string():
    lib = N.load()
    assert lib.psdr_abi_version() == N.ABI_VERSION
    assert isinstance(lib.psdr_last_error(), bytes)


def test_struct_sizes_match_header():
    # psdr_object: 14 int32; psdr_config: 14 uint32; psdr_camera: 11 floats + 3 int32
    assert ctypes.sizeof(N.PsdrObject) == 56
    assert ctypes.sizeof(N.PsdrConfig) == 56
    assert ctypes.sizeof(N.PsdrCamera) == 56


def test_argument_errors_without_gpu():
    lib = N.load()
    out = ctypes.c_void_p()
    assert lib.psdr_scene_create(None, ctypes.byref(out), None) == -1
    assert b"NULL" in lib.psdr_last_error()
    cam = N.PsdrCamera()
    assert lib.psdr_camera_matrices(ctypes.byref(cam), None, None, None, None, None) == -1  # height 0
    # more triangles than node child ids and 32-bit record offsets allow
    desc = N.PsdrSceneDesc()
    desc.n_tris = N.PSDR_MAX_TRIS + 1
    assert lib.psdr_scene_create(ctypes.byref(desc), ctypes.byref(out), None) == N.PSDR_ERR_ARG
    assert b"PSDR_MAX_TRIS" in lib.psdr_last_error()


def _cfg(**kw):
    secondary = kw.pop("secondary", (10000, 6, 6))
    base = dict(n_pass=1, spp_interior=16, enable_light_visable=0, spp_primary_edge=4, spp_secondary_edge=4,
                max_bounce=3, mis_light_samples=1, mis_bsdf_samples=1, primary_edge_preprocess_rounds=1,
                secondary_edge_preprocess_rounds=1, primary_edge_hypercube_resolution=10000)
    base.update(kw)
    c = N.PsdrConfig()
    for k, v in base.items():
        setattr(c, k, v)
    for i, r in enumerate(secondary):
        c.secondary_edge_hypercube_resolution[i] = r
    return c


def _cam(h, w):
    cam = N.PsdrCamera()
    cam.look_from[:] = (0.0, 1.0, 5.0)
    cam.up[:] = (0.0, 1.0, 0.0)
    cam.vfov, cam.height, cam.width, cam.kind = 0.785, h, w, N.PSDR_CAMERA_PERSPECTIVE
    return cam


@pytest.mark.parametrize("what,cfg,hw,msg", [
    # 2048^2 at 64 spp with one light and one BSDF sample: 268 M lanes x 3 rays x 3 floats > 2^31 - 1
    ("interior", dict(spp_interior=64), 2048, b"interior"),
    # 306 M lanes with 3 + 3 + 1 rays each (the examples' mis counts)
    ("interior-mis3", dict(spp_interior=73, mis_light_samples=3, mis_bsdf_samples=3), 2048, b"interior"),
    ("primary", dict(spp_primary_edge=64), 2048, b"primary"),
    ("primary-resolution", dict(primary_edge_hypercube_resolution=300_000_000), 64, b"primary"),
    ("direct-cells", dict(secondary=(100000, 100, 100)), 64, b"direct"),
    ("mis", dict(mis_light_samples=100000), 64, b"mis_light_samples"),
])
def test_render_rejects_sizes_whose_ray_indices_overflow(what, cfg, hw, msg):
    """Sizes whose 32-bit ray indices would overflow are refused with
    PSDR_ERR_ARG before any buffer or the scene is touched (the scene handle
    is NULL here, so a call that got past the check would report that)."""
    lib = N.load()
    c, cam = _cfg(**cfg), _cam(hw, hw)
    assert lib.psdr_render_forward(None, ctypes.byref(c), ctypes.byref(cam), None, 1, None) == N.PSDR_ERR_ARG
    err = lib.psdr_last_error()
    assert b"sizes out of range" in err and msg in err, err
    g = (ctypes.c_float * 1)()
    assert lib.psdr_render_backward(None, ctypes.byref(c), ctypes.byref(cam), g, None, None, None, None,
                                    N.PSDR_TERM_ALL, None) == N.PSDR_ERR_ARG
    assert b"sizes out of range" in lib.psdr_last_error()


def test_render_accepts_the_largest_configured_sizes():
    """The 1024^2 / 64 spp / depth-4 terrain render and the 512^2 / 16 spp step
    pass the size check (the NULL scene is reported instead)."""
    lib = N.load()
    for hw, spp in ((1024, 64), (512, 16)):
        c, cam = _cfg(spp_interior=spp), _cam(hw, hw)
        assert lib.psdr_render_forward(None, ctypes.byref(c), ctypes.byref(cam), None, 1, None) == N.PSDR_ERR_ARG
        assert b"sizes out of range" not in lib.psdr_last_error()


def test_tuning_roundtrip_and_validation(tmp_path):
    from torch_psdr_amd import tune
    assert ctypes.sizeof(N.PsdrTuning) == 24
    prev = tune.get_tuning()
    assert prev == tune.DEFAULTS
    try:
        tune.set_tuning(trace_blocks_per_cu=4, trace_stack_lds=32)
        assert tune.get_tuning() == dict(tune.DEFAULTS, trace_blocks_per_cu=4, trace_stack_lds=32)
        for bad in ({"trace_stack_lds": 12}, {"shade_blocks_per_cu": 0}, {"adjoint_blocks_per_cu": 65},
                    {"tile_lanes": 1000}, {"tile_lanes": -1}):
            with pytest.raises(RuntimeError):
                tune.set_tuning(**bad)
            # a rejected call changes nothing
            assert tune.get_tuning() == dict(tune.DEFAULTS, trace_blocks_per_cu=4, trace_stack_lds=32)
        with pytest.raises(KeyError):
            tune.set_tuning(stack=8)
        path = tmp_path / "tuning.json"
        tune.save("dev|w1", dict(tune.DEFAULTS, direct_blocks_per_cu=2), path)
        tune.save("dev|w2", dict(tune.DEFAULTS, shade_blocks_per_cu=16), path)
        assert tune.load("missing", path) is None
        got = tune.load("dev|w1", path)
        assert got == dict(tune.DEFAULTS, direct_blocks_per_cu=2) == tune.get_tuning()
        assert tune.load("dev|w2", path, apply=False)["shade_blocks_per_cu"] == 16
        assert tune.get_tuning()["shade_blocks_per_cu"] == tune.DEFAULTS["shade_blocks_per_cu"]
    finally:
        tune.set_tuning(prev)
