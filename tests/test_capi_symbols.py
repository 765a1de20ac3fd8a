"""The C-ABI library loads and exports every entry point include/sph_capi.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "sph_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 40
    assert "sph_ns_find_neighbors" in names and "sph_dfsph_step" in names


def test_library_exports_every_declared_symbol():
    from splishsplash_amd import capi
    L = capi.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_no_device_is_reported_not_faked():
    from splishsplash_amd import capi
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(capi.SPHError) as e:
        capi.TimeStepDFSPH()
    assert e.value.code == -5
    with pytest.raises(capi.SPHError):
        capi.NeighborhoodSearch(0.1)


def test_params_struct_matches_c_defaults():
    from splishsplash_amd import capi
    from splishsplash_amd.params import DFSPHParams, default_params
    p = DFSPHParams()
    capi.lib().sph_dfsph_default_params(ctypes.byref(p))
    q = default_params()
    for name, _ in DFSPHParams._fields_:
        a, b = getattr(p, name), getattr(q, name)
        if name in ("gravity", "reserved"):
            assert list(a) == list(b)
        else:
            assert a == b, name


def test_struct_layouts_agree_with_the_header():
    # The Python declarations and the oracle (compiled against the header)
    # must agree on the shared structs; a stale build or a field added on one
    # side only shows up here.
    from oracle import oracle
    from splishsplash_amd.params import DFSPHCounters, DFSPHParams
    L = oracle.lib()
    assert L.orc_sizeof_params() == ctypes.sizeof(DFSPHParams)
    assert L.orc_sizeof_counters() == ctypes.sizeof(DFSPHCounters)
    # a C compile of the header gives the same sizes and offsets
    src = os.path.join(ROOT, "include", "sph_capi.h")
    fields = [n for n, _ in DFSPHCounters._fields_ if n != "reserved"]
    code = "#include <stdio.h>\n#include <stddef.h>\n#include \"%s\"\nint main(void){printf(\"%%zu %%zu\", " \
           "sizeof(sph_dfsph_params_t), sizeof(sph_dfsph_counters_t));" % src
    code += "".join('printf(" %%zu", offsetof(sph_dfsph_counters_t, %s));' % f for f in fields) + "return 0;}\n"
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(code)
        subprocess.run(["gcc", "-o", exe, c], check=True)
        out = [int(v) for v in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == ctypes.sizeof(DFSPHParams) and out[1] == ctypes.sizeof(DFSPHCounters)
    assert out[2:] == [getattr(DFSPHCounters, f).offset for f in fields]
