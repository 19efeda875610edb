"""The C-ABI library loads and exports every entry point declared in include/smcm.h (no GPU
calls)."""
import os
import re
import subprocess

from resmcm_ari_amd import _build, smcm

HDR = os.path.join(_build.INCLUDE, "smcm.h")


def declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(smcm_[a-z_0-9]+)\s*\(", txt)))


def test_library_loads_and_exports_all():
    lib = smcm.load()
    names = declared(HDR)
    assert names, "no declarations parsed"
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(smcm.EXPORTS) == names


def test_dynamic_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", smcm.LIB_PATH], capture_output=True, text=True).stdout
    syms = set(re.findall(r" T (\w+)", out))
    for n in declared(HDR):
        assert n in syms, n
