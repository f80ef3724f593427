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
    assert ctypes.sizeof(kj.Opts) == 128
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


def test_option_checks_need_no_gpu():
    """Bad capacities are rejected before any device work."""
    lib = kj.load_library()
    X = (ctypes.c_float * 30)()
    m = kj.Matrix(ctypes.cast(X, ctypes.c_void_p), 10, 3, 3, 0)
    out_i = (ctypes.c_int32 * 30)()
    out_d = (ctypes.c_float * 30)()
    fake_ctx = ctypes.c_void_p(1)  # never dereferenced: the checks fail first
    for kw in ({"kcap": 96}, {"kcap": 2048}, {"kcap_mfma_max": 1024}, {"kcap_mfma_max": 48}):
        o = kj.make_opts(**kw)
        rc = lib.knnj_join(fake_ctx, m, m, 3, ctypes.byref(o), out_i, out_d, None)
        assert rc == kj.KNNJ_EINVAL, kw
        assert b"kcap" in lib.knnj_last_error()


def test_plan_part_table_matches_header():
    text = open(os.path.join(ROOT, "include", "knnj.h")).read()
    ids = dict((n.lower(), int(v)) for n, v in re.findall(r"KNNJ_PART_(\w+)\s*=\s*(\d+)", text))
    assert {n: v[0] for n, v in kj.PLAN_PARTS.items()} == {
        "r_reps": ids["r_reps"], "r_radius": ids["r_radius"], "r_offsets": ids["r_offsets"],
        "r_perm": ids["r_perm"], "r_pdist": ids["r_pdist"], "q_reps": ids["q_reps"],
        "q_radius": ids["q_radius"], "q_offsets": ids["q_offsets"], "q_perm": ids["q_perm"],
        "q_pdist": ids["q_pdist"], "theta": ids["theta"], "pair_count": ids["pair_count"],
        "pair_list": ids["pair_list"], "pair_lb": ids["pair_lb"]}


def test_null_context_is_rejected():
    lib = kj.load_library()
    m = kj.Matrix(None, 0, 3, 3, 0)
    rc = lib.knnj_join(None, m, m, 1, None, None, None, None)
    assert rc == kj.KNNJ_EINVAL
    assert b"context" in lib.knnj_last_error()


def test_struct_layouts_match_c_compiler(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler
    gives the header's structs."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    mirrors = {"knnj_matrix": kj.Matrix, "knnj_opts": kj.Opts, "knnj_stats": kj.Stats,
               "knnj_instance": kj.Instance, "knnj_index_desc": kj.IndexDesc,
               "knnj_server_config": kj.ServerConfig, "knnj_request_info": kj.RequestInfo,
               "knnj_plan_desc": kj.PlanDesc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "knnj.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([gcc, f"-I{inc}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            cname, fname, v = ln.split()
            got[(cname, fname)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
