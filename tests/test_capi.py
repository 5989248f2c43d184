"""The C-ABI library loads and exports every symbol include/fricc_capi.h
declares (CPU only: no compute calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "fricc_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fricc_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from fricc import _capi

    L = _capi.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_capi.EXPORTS)
    assert L.fricc_abi_version() == 1


def test_library_is_built_for_gfx950():
    from fricc import _capi

    data = open(_capi.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_oracle_is_separate_from_product():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "fri-cc_amd", "fricc")
    for f in os.listdir(pkg):
        if f.endswith(".py"):
            txt = open(os.path.join(pkg, f)).read()
            assert "oracle" not in txt.replace("oracle)", "").lower() or "import oracle" not in txt
            assert "from oracle" not in txt and "import oracle" not in txt


def test_last_error_is_a_string():
    from fricc import _capi

    assert isinstance(_capi.lib().fricc_last_error(), bytes)
