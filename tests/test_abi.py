"""CPU checks of the C ABI: the library loads and exports every entry point
that include/knnj.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import pytest

import knn_sweet_amd as kj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "knnj.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(knnj_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = kj.load_library()
    names = declared_functions()
    assert "knnj_join" in names and "knnj_ctx_create" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(kj.EXPORTS) <= set(names)


def test_struct_layouts_match_header():
    # field order and sizes of the ctypes mirrors (x86-64 ABI)
    assert ctypes.sizeof(kj.Matrix) == 40
    assert kj.Opts.seed.offset == 16
    assert ctypes.sizeof(kj.Opts) == 16 + 8 + 4 * 4 + 6 * 4
    assert kj.Stats.n_dist_evals.offset == 40


def test_version_and_error_string():
    assert kj.version().startswith("knnj")
    assert isinstance(kj.load_library().knnj_last_error(), bytes)


def test_context_without_gpu_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(kj.KnnjError) as e:
        kj.Context(0)
    assert e.value.status == kj.KNNJ_EDEVICE


def test_null_context_is_rejected():
    lib = kj.load_library()
    m = kj.Matrix(None, 0, 3, 3, 0)
    rc = lib.knnj_join(None, m, m, 1, None, None, None, None)
    assert rc == kj.KNNJ_EINVAL
    assert b"context" in lib.knnj_last_error()
