"""The C-ABI library loads and exports every symbol include/swx.h declares
(no compute calls: this runs on machines without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import REPO

HEADER = os.path.join(REPO, "include", "swx.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(swx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "swx_plan_create" in syms and "swx_adjoint" in syms
    from seismicwaves import _native
    assert set(syms) == set(_native.EXPORTS)


def test_library_exports_every_declared_symbol():
    from seismicwaves import _native
    if not os.path.exists(_native.LIBPATH):
        pytest.skip("libswx.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_native.LIBPATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    lib.swx_abi_version.restype = ctypes.c_int
    assert lib.swx_abi_version() == 1


def test_struct_layouts_match_header():
    """ctypes mirrors of the header's structs have the sizes (and offsets)
    the C compiler gives them."""
    import subprocess
    import tempfile
    from seismicwaves import _native
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "swx.h"
int main(void){printf("%zu %zu %zu %zu %zu %zu\n", sizeof(swx_plan_desc), sizeof(swx_shot_desc),
  sizeof(swx_plan_info), sizeof(swx_taps), sizeof(swx_elastic_shot_desc),
  offsetof(swx_elastic_shot_desc, rec));return 0;}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == [ctypes.sizeof(_native.PlanDesc), ctypes.sizeof(_native.ShotDesc),
                                     ctypes.sizeof(_native.PlanInfo), ctypes.sizeof(_native.Taps),
                                     ctypes.sizeof(_native.ElasticShotDesc), _native.ElasticShotDesc.rec.offset]
