"""The C ABI: libwpbr_hip.so loads and exports every entry point include/wpbr.h
declares (no GPU needed; no compute call is made)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from woodenpbrengine_amd import LIB_PATH, lib

ROOT = Path(__file__).resolve().parent.parent


def declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(wpbr_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    assert "wpbr_lbvh_build" in names and "wpbr_collide" in names
    assert set(lib.EXPORTS) <= set(names)


def test_library_exports_every_declared_symbol():
    assert LIB_PATH.exists(), "run __graft_entry__.build() first"
    so = C.CDLL(str(LIB_PATH))
    missing = [n for n in declared() if not hasattr(so, n)]
    assert not missing, missing


def test_status_strings_without_gpu():
    so = lib.load()
    assert so.wpbr_status_string(0) == b"ok"
    assert so.wpbr_status_string(4) == b"traversal stack overflow"


def test_invalid_arguments_rejected_without_gpu():
    so = lib.load()
    assert so.wpbr_open(0, None) == 1
    assert so.wpbr_set_trace_config(None, None) == 1
