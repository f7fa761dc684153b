"""The C-ABI library loads on the CPU and exports every symbol include/jztree_amd.h declares."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "jztree_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(jzt_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(os.path.join(ROOT, "jztree_amd", "libjztree_amd.so"))
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_error_string():
    import jztree_amd._lib as L
    assert L.LIB.jzt_abi_version() == 1
    assert isinstance(L.LIB.jzt_last_error(), bytes)


def test_tmp_size_queries_are_host_only():
    import jztree_amd._lib as L
    assert L.LIB.jzt_zsort_tmp_bytes(1 << 20, 3, 0) >= 24 * (1 << 20)
    # 256 bytes of work-queue counters after the multi-pass state (aligned to 256 bytes): none for
    # one pass, 8 bytes per query past k = 64 for float32, 12 past k = 32 for float64
    assert L.LIB.jzt_knn_leaf2leaf_tmp_bytes(1000, 16, 0) == 256
    assert L.LIB.jzt_knn_leaf2leaf_tmp_bytes(1000, 100, 0) == 8192 + 256
    assert L.LIB.jzt_knn_leaf2leaf_tmp_bytes(1000, 40, 1) == 12032 + 256


def _param_kinds():
    """Per declared function: one kind per parameter, 'p' pointer, 'i'/'f' integer/float scalar."""
    src = open(os.path.join(ROOT, "include", "jztree_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(jzt_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", src):
        kinds = []
        for p in m.group(2).split(","):
            p = p.strip()
            if not p or p == "void":
                continue
            if "*" in p or "jzt_stream_t" in p:
                kinds.append("p")
            elif p.startswith("float") or p.startswith("double"):
                kinds.append("f")
            else:
                kinds.append("i")
        out[m.group(1)] = kinds
    return out


def test_ctypes_signatures_match_header():
    """Every ctypes binding passes exactly the header's parameters, pointer/int/float kinds in order."""
    import ctypes as C
    from jztree_amd._lib import _SIGNATURES
    kinds = _param_kinds()
    def kind(t):
        if t is C.c_void_p or t is C.c_char_p:
            return "p"
        if t in (C.c_float, C.c_double):
            return "f"
        return "i"
    assert set(_SIGNATURES) == set(kinds)
    for name, (args, _) in _SIGNATURES.items():
        assert [kind(a) for a in args] == kinds[name], name
