"""The C-ABI library loads on a machine without a GPU and exports every entry
point include/panohip.h declares (no device work is started here)."""
import os
import re

import panostitcher_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(pano_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = P.lib()
    decl = declared_symbols()
    assert len(decl) >= 25
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(P.exported_symbols()) == decl


def test_status_strings_and_host_helpers():
    assert P.status_string(0) == "ok"
    assert P.status_string(-6) == "device memory budget exceeded"
    assert P.lib().pano_version() >= 100
    lut = P.get_lut(0.5)
    assert lut[3] == 2 and lut[5] == 2 and lut[255] == 128


def test_estimate_memory_is_monotone():
    e1 = P.CudaTilingMultibandBlendFast.estimate_memory_size(2048, 1024, 2)
    e2 = P.CudaTilingMultibandBlendFast.estimate_memory_size(2048, 1024, 4)
    e3 = P.CudaTilingMultibandBlendFast.estimate_memory_size(4096, 2048, 4)
    assert 0 < e1 < e2 < e3
