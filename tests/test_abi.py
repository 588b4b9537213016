"""The C ABI library loads and exports every symbol that include/*.h declares (no GPU
calls), and the engine refuses to run without a GPU instead of falling back."""
import ctypes
import pathlib
import re

import pytest

import gems_amd as G

INC = pathlib.Path(__file__).resolve().parent.parent / "include"


def declared():
    names = set()
    for h in INC.glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(plk_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_all_declared_symbols_exported():
    lib = G.lib()
    names = declared()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_headers():
    assert declared() <= set(G.declared_symbols()) | {"plk_tree_context"} | set(G.declared_symbols())


def test_memory_required_formula():
    m = G.lib().plk_memory_required(20, 4, 100, 50000, 98, 0, 0)
    # partials dominate: 98 slots x 50048 patterns x 80 x 8 bytes
    assert m > 98 * 50048 * 80 * 8
    assert m < 1.2 * 98 * 50048 * 80 * 8


@pytest.mark.skipif(G.device_count() > 0, reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    h = ctypes.c_void_p()
    rc = G.lib().plk_create(ctypes.byref(h), 4, 4, 8, 100, 6, 0, 0)
    assert rc != 0
    assert b"no HIP device" in G.lib().plk_last_error()
