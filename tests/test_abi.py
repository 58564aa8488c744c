"""The C ABI library loads on a machine without a GPU and exports every
function include/*.h declares; the Python mirror binds them all.  No compute
call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "plan_amd", "libplan_hip.so")


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b(plan_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("plan_ctx_create", "plan_build_tree", "plan_find_planetesimals", "plan_knn_density", "plan_hop"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_mirror_binds_all():
    import plan_amd

    lib = plan_amd.lib()
    for n in declared_functions():
        assert getattr(lib, n) is not None


def test_missing_library_fails_loudly(monkeypatch):
    import plan_amd

    monkeypatch.setattr(plan_amd, "_lib", None)
    monkeypatch.setattr(plan_amd, "LIB_PATH", os.path.join(ROOT, "plan_amd", "no_such_lib.so"))
    with pytest.raises(plan_amd.PlanError):
        plan_amd.lib()
