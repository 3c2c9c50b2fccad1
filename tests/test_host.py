"""Host-side checks that need no GPU: the C-ABI library loads and exports every entry
point include/tbfm_gpu.h declares, input files round-trip, the CLI and bench contract."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import cturbobfm_amd
from cturbobfm_amd import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "tbfm_gpu.h")).read()
    return sorted(set(re.findall(r"\b(tbfm_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(cturbobfm_amd.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(cturbobfm_amd.EXPORTS) == syms


def test_creation_without_device_fails_cleanly():
    try:
        import torch  # noqa: F401
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(cturbobfm_amd.SolverError):
        cturbobfm_amd.Solver.from_case(cases.bump2d(9, 6))


def test_case_files_round_trip(tmp_path):
    case = cases.rotor3d(9, 5, 7)
    ini = cases.write_case(case, str(tmp_path / "rotor"))
    back = cases.read_case(ini)
    assert (back["ni"], back["nj"], back["nk"]) == (9, 5, 7)
    for k in ("x", "y", "z"):
        assert np.array_equal(back[k], case[k])
    for f in case["fields"]:
        assert np.array_equal(back["fields"][f], case["fields"][f])
    assert [tuple(p) for p in back["patches"]] == [tuple(p) for p in case["patches"]]


def test_full_annulus_closure_is_bitwise():
    case = cases.rotor3d(9, 5, 7)
    assert np.array_equal(case["y"][:, :, 0], case["y"][:, :, -1])
    assert np.array_equal(case["z"][:, :, 0], case["z"][:, :, -1])


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_cli_usage_message():
    cli = os.path.join(ROOT, "cturbobfm_amd", "turbobfm")
    r = subprocess.run([cli], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr
