"""CPU checks of the drop-in boundary: the library loads without a GPU and
exports every function that include/kssdhip.h declares (no compute calls)."""
import ctypes
import re

import kssdtest as kt


def declared_functions():
    text = (kt.ROOT / "include" / "kssdhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(kssdhip_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("kssdhip_index_open", "kssdhip_count_block", "kssdhip_pair_features",
                 "kssdhip_find_first", "kssdhip_index_close", "kssdhip_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(kt.PKG / "libkssdhip.so"))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_without_device_are_reported_not_fatal():
    lib = kt.load_kssdhip()
    h = ctypes.c_void_p()
    # argument validation happens before any device work
    assert lib.kssdhip_index_open(None, 5, 1, 20, ctypes.byref(h)) == -1
    assert b"null records" in lib.kssdhip_last_error()
    assert lib.kssdhip_index_open(None, 0, 0, 20, ctypes.byref(h)) == -1
    assert lib.kssdhip_count_block(None, None, None, 1, 0, 0, None, None, None, None) == -1
