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


def test_tuning_roundtrip_and_validation(tmp_path):
    from torch_psdr_amd import tune
    assert ctypes.sizeof(N.PsdrTuning) == 20
    prev = tune.get_tuning()
    assert prev == tune.DEFAULTS
    try:
        tune.set_tuning(trace_blocks_per_cu=4, trace_stack_lds=32)
        assert tune.get_tuning() == dict(tune.DEFAULTS, trace_blocks_per_cu=4, trace_stack_lds=32)
        for bad in ({"trace_stack_lds": 12}, {"shade_blocks_per_cu": 0}, {"adjoint_blocks_per_cu": 65}):
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
