"""The drop-in structs of include/fame_main.h have the reference's binary layout
(include/FAME_Internal_Common.h:29-211 of the reference tree, tabulated in
tests/golden/par_layout.json by tests/golden/make_par_layout.py), in C and in
C++, and the library exports FAME_Main_Code under the reference's C++ symbol
(include/FAME_Main_Code.cuh:4) as well as a C symbol of the same name."""
import ctypes as C
import json
import os
import shutil
import sys

import pytest

import fame_amd as F

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_par_layout as ML  # noqa: E402
import par_structs as PS  # noqa: E402


def _golden():
    with open(os.path.join(HERE, "golden", "par_layout.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("compiler,flags", [("gcc", ["-std=c11"]), ("g++", ["-std=c++17"])])
def test_header_layout_matches_reference(tmp_path, compiler, flags):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not available")
    ours = ML.run_layout(compiler, os.path.join(ROOT, "include"), '#include "fame_main.h"', flags, str(tmp_path))
    ref = _golden()["layout"]
    assert ours == ref


def test_ctypes_mirror_matches_reference():
    ref = _golden()["layout"]
    for st, members in ML.FIELDS:
        cls = getattr(PS, st)
        assert C.sizeof(cls) == ref[st]["sizeof"], st
        for m in members:
            assert getattr(cls, m).offset == ref[st][m], (st, m)


def test_main_code_symbols():
    sym = _golden()["FAME_Main_Code_symbol"]
    assert sym == "_Z14FAME_Main_Code3PARP7PROFILE"
    lib = C.CDLL(F.LIB_PATH)
    assert hasattr(lib, sym)              # what a C++ caller of the reference links against
    assert hasattr(lib, "FAME_Main_Code")  # C / FFI alias
    assert hasattr(lib, "fame_main_code_run")
