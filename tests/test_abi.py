"""The C-ABI libraries load and export every entry point their headers declare (no
compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(rf"\b({prefix}_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,prefix,lib", [
    ("chronos_amg.h", "chr", "chronoslab_amd/libchronos_amg.so"),
    ("chronos_setup.h", "chs", "chronoslab_amd/libchronos_setup.so"),
])
def test_exports(header, prefix, lib):
    path = os.path.join(ROOT, lib)
    if not os.path.exists(path):
        if "amg" in lib:
            from chronoslab_amd import build as B

            B.build_amg()
        else:
            from chronoslab_amd import build as B

            B.build_setup()
    so = ctypes.CDLL(path)
    names = _declared(header, prefix)
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from chronoslab_amd import _amg_sigs

    assert set(_declared("chronos_amg.h", "chr")) == set(_amg_sigs.EXPORTS)


def test_oracle_is_not_imported_by_the_package():
    for dirpath, _, files in os.walk(os.path.join(ROOT, "chronoslab_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "oracle." not in src.replace("oracle.h", ""), f
