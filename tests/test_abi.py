"""The C-ABI library loads and exports every function include/lbm_c_api.h declares
(no compute calls: runs without a GPU)."""
import os
import re

import lbm
from conftest import ROOT


def _declared():
    text = open(os.path.join(ROOT, "include", "lbm_c_api.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = lbm.lib()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    bound = {n for n, _, _ in lbm.SIGNATURES}
    assert set(_declared()) == bound


def test_errors_are_reported_through_the_abi():
    import pytest

    with pytest.raises(lbm.LbmError) as e:
        lbm.Geometry.from_config("/nonexistent/config.txt")
    assert e.value.code == lbm.ERR_IO
    assert "Unable to open" in str(e.value)
