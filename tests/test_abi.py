"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import os
import re

import pytest

import tortoise_amd as tg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "tortoise_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tg_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert _declared() == sorted(tg.EXPORTS)


def test_library_exports_every_symbol():
    if not os.path.exists(tg.library_path()):
        pytest.fail("libtortoise_gpu.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(tg.library_path())
    for name in _declared():
        assert hasattr(lib, name), name


def test_num_nodes_without_device():
    L = tg.load()
    ns = (ctypes.c_int32 * 3)(32, 32, 1)
    assert L.tg_num_nodes(2, ns) == 1089
    ns3 = (ctypes.c_int32 * 3)(20, 20, 20)
    assert L.tg_num_nodes(3, ns3) == 9261
    assert L.tg_num_nodes(4, ns3) == -1


def test_bad_descriptor_is_rejected_without_device():
    L = tg.load()
    d = tg.ProblemDesc()
    d.dim = 5
    h = ctypes.c_void_p()
    assert L.tg_ctx_create(ctypes.byref(d), ctypes.byref(h)) == tg.TG_ERR_ARG
    assert not h.value
