"""The C ABI library loads and exports every function include/nadia_amd.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nadia_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(nd):
    names = declared(os.path.join(ROOT, "include", "nadia_amd.h"))
    assert len(names) > 40
    lib = ctypes.CDLL(nd.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_only_entry_points(nd):
    lib = nd.library()
    assert lib.nadia_version() >= 100
    for n, ok, native in ((2048, 1, 1), (384, 1, 1), (768, 1, 1), (24, 1, 1), (100, 1, 0), (6, 1, 0), (1, 1, 0),
                          (4096, 1, 1), (4097, 0, 0), (8192, 0, 0), (0, 0, 0)):
        assert lib.nadia_size_supported(n) == ok
        assert lib.nadia_size_native(n) == native
    assert nd.get_alg_from_name("RAAR") == nd.RAAR
    assert nd.get_alg_from_name("nope") == -1
