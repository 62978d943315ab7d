"""The C-ABI library loads without a GPU and exports every symbol that
include/mtimager_amd.h, include/mtimager_amd_fits.h and
include/mtimager_amd_ingest.h declare (no compute
calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("mtimager_amd.h", "mtimager_amd_fits.h", "mtimager_amd_ingest.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(mti_[a-z0-9_]+)\s*\(", src))
    return names


def test_header_and_binding_agree(mti):
    from mtimager_amd import ingest
    assert _declared() == set(mti.EXPORTS) | set(ingest.EXPORTS)


def test_library_exports_every_declared_symbol(mti):
    L = ctypes.CDLL(mti.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(L, s)]
    assert not missing, missing


def test_version_string(mti):
    assert mti.lib().mti_version().decode().startswith("mtimager_amd")


def test_no_cpu_fallback_in_product_path():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "mt-imager_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                text = open(os.path.join(dp, f)).read()
                assert "import oracle" not in text and "mti_oracle" not in text, f
