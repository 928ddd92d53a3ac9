"""CPU checks of the drop-in boundary: the C ABI library loads and exports every function that
include/mbxg.h declares (no device calls are made)."""
import os
import re

import pytest

from conftest import ROOT

from mbx_amd import engine


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mbxg_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libmbxg.so not built")
    lib = engine.lib()
    names = declared("mbxg.h")
    assert names, "no declarations found"
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding prototypes every declared function
    assert set(names) <= set(engine.PROTOTYPES)


def test_last_error_without_device():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libmbxg.so not built")
    assert isinstance(engine.lib().mbxg_last_error(), bytes)
