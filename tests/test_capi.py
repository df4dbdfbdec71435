"""The C-ABI library loads and exports every entry point of include/mmc.h;
host-side argument checking (no device work)."""
import ctypes as C
import re

from conftest import ROOT


def _declared():
    text = "".join(p.read_text() for p in sorted((ROOT / "include").glob("*.h")))
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(mmc_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree(mmc):
    assert _declared() == sorted(mmc.EXPORTS)


def test_library_exports_every_symbol(mmc):
    L = C.CDLL(mmc.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(L, s)]
    assert not missing, missing


def test_null_and_bad_arguments_are_rejected(mmc):
    L = mmc.lib()
    out = C.c_void_p()
    assert L.mmc_create(None, 0, C.byref(out)) != 0
    p = mmc.make_params()
    assert L.mmc_create(C.byref(p), 7, C.byref(out)) != 0      # unknown precision
    assert L.mmc_batch_create(C.byref(p), 0, 0, C.byref(out)) != 0
    assert L.mmc_last_error(None) == b"null context"
    assert L.mmc_count(None) == 0
    assert L.mmc_sweeps(None, 0, 0, 1) != 0
    assert L.mmc_set_option(None, b"graph", 1) != 0
